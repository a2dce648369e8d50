//#####################################################################
// Drop-in of the MI355X explicit APIC step inside PhysBAM itself: the adapter
// a maintainer adds beside MPM_DRIVER (compiled against PhysBAM's own
// headers, linked with libmpmx.so; it is not part of libmpmx).
//
// MPM_GPU_STEP<TV> reads the scene from an MPM_EXAMPLE<TV> -- grid,
// MPM_PARTICLES, the one MPM_FINITE_ELEMENTS force and its constitutive
// model, MPM_GRAVITY, box collision objects, clamped or Drucker-Prager
// plasticity, reflection sides and the strong/sound-speed CFL flags
// (Hybrid_Methods/Examples_And_Drivers/MPM_EXAMPLE.h:31-176) -- refuses what
// the GPU step does not do with PHYSBAM_FATAL_ERROR, and replaces
// MPM_DRIVER::Advance_One_Time_Step (MPM_DRIVER.cpp:141-157).  Callbacks the
// example registered with Add_Callbacks (MPM_EXAMPLE.cpp:350-356) on the
// named steps fire around each of them, as MPM_DRIVER::Step fires them
// (MPM_DRIVER.cpp:1380-1393); such a step runs stage by stage.
//
// Refused: moving collision objects (func_frame / func_twist,
// Hybrid_Methods/Collisions/MPM_COLLISION_OBJECT.h:23-24), implicit objects
// other than ANALYTIC_IMPLICIT_OBJECT<RANGE<TV> >, weights other than the
// quadratic spline, midpoint / lag_Dp / Oldroyd / FLIP variants, the implicit
// step, Lagrangian forces and more than one elastic force.
//#####################################################################
#pragma once
#include <Core/Log/LOG.h>
#include <Deformables/Constitutive_Models/COROTATED_FIXED.h>
#include <Deformables/Constitutive_Models/NEO_HOOKEAN.h>
#include <Deformables/Constitutive_Models/ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.h>
#include <Geometry/Implicit_Objects/ANALYTIC_IMPLICIT_OBJECT.h>
#include <Hybrid_Methods/Collisions/MPM_COLLISION_OBJECT.h>
#include <Hybrid_Methods/Examples_And_Drivers/MPM_EXAMPLE.h>
#include <Hybrid_Methods/Examples_And_Drivers/MPM_PARTICLES.h>
#include <Hybrid_Methods/Forces/MPM_DRUCKER_PRAGER.h>
#include <Hybrid_Methods/Forces/MPM_FINITE_ELEMENTS.h>
#include <Hybrid_Methods/Forces/MPM_GRAVITY.h>
#include <Hybrid_Methods/Forces/MPM_PLASTICITY_CLAMP.h>
#include <Hybrid_Methods/Iterators/PARTICLE_GRID_WEIGHTS_SPLINE.h>
#include <vector>

#include "mpmx.h"

namespace PhysBAM{

template<class TV>
class MPM_GPU_STEP
{
    typedef typename TV::SCALAR T;
    enum {d=TV::m};
public:
    MPM_EXAMPLE<TV>& ex;
    mpmx_ctx* ctx=0;

    static void Check(int rc){if(rc!=MPMX_OK) PHYSBAM_FATAL_ERROR(std::string("mpmx: ")+mpmx_last_error());}

    // The scene is checked in full before the device is touched.
    explicit MPM_GPU_STEP(MPM_EXAMPLE<TV>& example)
        :ex(example)
    {
        if(!ex.use_affine || !ex.use_symplectic_euler || ex.flip!=0 || ex.use_midpoint || ex.lag_Dp || ex.use_oldroyd)
            PHYSBAM_FATAL_ERROR("the GPU step is the explicit symplectic APIC step (flip 0, no midpoint, lag_Dp or Oldroyd)");
        if(!dynamic_cast<PARTICLE_GRID_WEIGHTS_SPLINE<TV,2>*>(ex.weights))
            PHYSBAM_FATAL_ERROR("the GPU step uses quadratic B-spline weights (Set_Weights(2))");
        if(ex.ghost!=3) PHYSBAM_FATAL_ERROR("the GPU step uses three ghost layers");
        if(ex.lagrangian_forces.m) PHYSBAM_FATAL_ERROR("Lagrangian forces are not part of the GPU step");
        int model=-1;double threshold=0.25;
        double gravity[3]={0,0,0};
        for(int i=0;i<ex.forces.m;i++){
            if(auto* fe=dynamic_cast<MPM_FINITE_ELEMENTS<TV>*>(ex.forces(i))){
                if(model>=0) PHYSBAM_FATAL_ERROR("one MPM_FINITE_ELEMENTS force (per-particle mu, lambda) on the GPU step");
                if(dynamic_cast<COROTATED_FIXED<T,d>*>(&fe->constitutive_model)) model=MPMX_FIXED_COROTATED;
                else if(auto* nh=dynamic_cast<NEO_HOOKEAN<T,d>*>(&fe->constitutive_model)){
                    model=MPMX_NEO_HOOKEAN;threshold=nh->failure_threshold;}
                else if(dynamic_cast<ST_VENANT_KIRCHHOFF_HENCKY_STRAIN<T,d>*>(&fe->constitutive_model)) model=MPMX_HENCKY;
                else PHYSBAM_FATAL_ERROR("constitutive model not on the GPU step");}
            else if(auto* gr=dynamic_cast<MPM_GRAVITY<TV>*>(ex.forces(i))){
                for(int a=0;a<d;a++) gravity[a]+=gr->gravity(a);}
            else PHYSBAM_FATAL_ERROR("force not on the GPU step");}
        std::vector<const ANALYTIC_IMPLICIT_OBJECT<RANGE<TV> >*> boxes;
        for(int i=0;i<ex.collision_objects.m;i++){
            const MPM_COLLISION_OBJECT<TV>* co=ex.collision_objects(i);
            if(co->func_frame || co->func_twist)
                PHYSBAM_FATAL_ERROR("moving collision objects (func_frame, func_twist) are not on the GPU step");
            auto* box=dynamic_cast<const ANALYTIC_IMPLICIT_OBJECT<RANGE<TV> >*>(co->io);
            if(!box) PHYSBAM_FATAL_ERROR("the GPU step takes box collision objects (ANALYTIC_IMPLICIT_OBJECT<RANGE<TV> >)");
            boxes.push_back(box);}
        const MPM_PLASTICITY_CLAMP<TV>* clamp=0;
        const MPM_DRUCKER_PRAGER<TV>* dp=0;
        T cohesion=0;
        for(int i=0;i<ex.plasticity_models.m;i++){
            if(i) PHYSBAM_FATAL_ERROR("one plasticity model on the GPU step");
            clamp=dynamic_cast<const MPM_PLASTICITY_CLAMP<TV>*>(ex.plasticity_models(i));
            dp=dynamic_cast<const MPM_DRUCKER_PRAGER<TV>*>(ex.plasticity_models(i));
            if(!clamp && !dp) PHYSBAM_FATAL_ERROR("plasticity model not on the GPU step");
            if(dp){
                cohesion=dp->sigma_Y.m?dp->sigma_Y(0):0;
                for(int p=0;p<dp->sigma_Y.m;p++)
                    if(dp->sigma_Y(p)!=cohesion) PHYSBAM_FATAL_ERROR("the GPU step takes one Drucker-Prager cohesion");}}

        mpmx_grid_desc g={d,(int)sizeof(T),{1,1,1},{0,0,0},{1,1,1}};
        for(int a=0;a<d;a++){
            g.cells[a]=ex.grid.counts(a);
            g.lo[a]=ex.grid.domain.min_corner(a);
            g.hi[a]=ex.grid.domain.max_corner(a);}
        Check(mpmx_create(&g,&ctx));
        Check(mpmx_set_constitutive(ctx,model<0?MPMX_FIXED_COROTATED:model,threshold));
        Check(mpmx_set_gravity(ctx,gravity));
        for(size_t i=0;i<boxes.size();i++){
            const MPM_COLLISION_OBJECT<TV>* co=ex.collision_objects(i);
            double lo[3]={0,0,0},hi[3]={0,0,0};
            for(int a=0;a<d;a++){lo[a]=boxes[i]->analytic.min_corner(a);hi[a]=boxes[i]->analytic.max_corner(a);}
            Check(mpmx_add_collision_box(ctx,(int)co->type,(double)co->friction,lo,hi));}
        if(ex.reflection_bc){
            int sides[6]={0,0,0,0,0,0};
            for(int i=0;i<2*d;i++) sides[i]=(int)ex.side_bc_type(i);
            Check(mpmx_set_reflection(ctx,1,ex.use_full_reflection,(double)ex.reflection_bc_friction,sides));}
        if(ex.use_strong_cfl || ex.use_sound_speed_cfl)
            Check(mpmx_set_dt_limits(ctx,ex.use_strong_cfl,ex.use_sound_speed_cfl,(double)ex.cfl,(double)ex.cfl_F,
                (double)ex.cfl_sound,(double)ex.min_dt));
        Upload();
        if(clamp) Check(mpmx_set_plasticity(ctx,1,clamp->theta_c,clamp->theta_s,clamp->max_hardening,clamp->hardening_factor));
        if(dp) Check(mpmx_set_drucker_prager(ctx,1,dp->a0,dp->a1,dp->a3,dp->a4,cohesion,dp->use_vc));
    }
    ~MPM_GPU_STEP(){mpmx_destroy(ctx);}

    // example.particles -> device (MPM_PARTICLES X, V, F, B, mass, volume, mu, lambda, valid)
    void Upload()
    {
        MPM_PARTICLES<TV>& P=ex.particles;
        std::vector<unsigned char> v(P.number);
        for(int p=0;p<P.number;p++) v[p]=P.valid(p);
        Check(mpmx_set_particles(ctx,P.number,P.X.Get_Array_Pointer(),P.V.Get_Array_Pointer(),P.F.Get_Array_Pointer(),
            P.B.Get_Array_Pointer(),P.mass.Get_Array_Pointer(),P.volume.Get_Array_Pointer(),P.mu.Get_Array_Pointer(),
            P.lambda.Get_Array_Pointer(),v.data()));
    }
    // device -> example.particles (X, V, F, B, valid)
    void Download()
    {
        MPM_PARTICLES<TV>& P=ex.particles;
        std::vector<unsigned char> v(P.number);
        Check(mpmx_get_particles(ctx,P.X.Get_Array_Pointer(),P.V.Get_Array_Pointer(),P.F.Get_Array_Pointer(),
            P.B.Get_Array_Pointer(),v.data()));
        for(int p=0;p<P.number;p++) P.valid(p)=v[p]!=0;
    }
    // device -> example.mass, example.velocity over Domain_Indices(ghost)
    void Download_Grid()
    {
        Check(mpmx_get_grid(ctx,ex.mass.array.Get_Array_Pointer(),ex.velocity.array.Get_Array_Pointer()));
    }

    // MPM_DRIVER::Advance_One_Time_Step with example.dt; Reduce_Dt may lower
    // example.dt.  With callbacks on any named step, each runs between them.
    void Advance_One_Time_Step()
    {
        bool staged=false;
        for(int k=0;k<MPMX_STAGES && !staged;k++){
            const auto* p=ex.time_step_callbacks.Get_Pointer(mpmx_stage_name(k));
            staged=p && (p->y(0).m || p->y(1).m);}
        if(staged){
            for(int k=0;k<MPMX_STAGES;k++){
                const char* name=mpmx_stage_name(k);
                Fire(name,true);
                Check(mpmx_step_stage(ctx,(double)ex.dt,k));
                Fire(name,false);}}
        else Check(mpmx_step(ctx,(double)ex.dt,1));
        double taken=0;
        Check(mpmx_last_dt(ctx,&taken));
        ex.dt=(T)taken;
    }

    // MPM_DRIVER::Compute_Dt (MPM_DRIVER.cpp:1098-1104) of the device particles
    T Compute_Dt()
    {
        double dt=0;
        Check(mpmx_compute_dt(ctx,(double)ex.cfl,(double)ex.max_dt,&dt));
        return (T)dt;
    }

private:
    // the begin (y(1)) or end (y(0)) callbacks of a named step, as MPM_DRIVER::Step runs them
    void Fire(const char* name,bool begin)
    {
        const auto* p=ex.time_step_callbacks.Get_Pointer(name);
        if(!p) return;
        const auto& list=p->y(begin?1:0);
        for(int i=0;i<list.m;i++) list(i)();
    }
};
}
