//#####################################################################
// Scene file reader and the reference MPM_EXAMPLE / MPM_DRIVER set-up shared
// by the reference harness (ref_harness.cpp) and the drop-in check
// (adapter_check.cpp).  Test infrastructure, compiled only against the
// reference checkout; the scene layout is written by tests/golden/scene_io.py.
//#####################################################################
#pragma once
#include <Core/Log/LOG.h>
#include <Core/Math_Tools/clamp.h>
#include <Core/Math_Tools/RANGE_ITERATOR.h>
#include <Core/Matrices/DIAGONAL_MATRIX.h>
#include <Core/Matrices/MATRIX.h>
#include <Grid_Tools/Grids/GRID.h>
#include <Deformables/Constitutive_Models/COROTATED_FIXED.h>
#include <Deformables/Constitutive_Models/NEO_HOOKEAN.h>
#include <Hybrid_Methods/Examples_And_Drivers/MPM_DRIVER.h>
#include <Core/Read_Write/FILE_UTILITIES.h>
#include <Core/Random_Numbers/RANDOM_NUMBERS.h>
#include <Geometry/Basic_Geometry/SPHERE.h>
#include <Geometry/Implicit_Objects/ANALYTIC_IMPLICIT_OBJECT.h>
#include <Geometry/Seeding/POISSON_DISK.h>
#include <Hybrid_Methods/Examples_And_Drivers/MPM_EXAMPLE.h>
#include <Hybrid_Methods/Examples_And_Drivers/MPM_PARTICLES.h>
#include <Hybrid_Methods/Forces/MPM_FINITE_ELEMENTS.h>
#include <Hybrid_Methods/Forces/MPM_GRAVITY.h>
#include <Hybrid_Methods/Forces/MPM_PLASTICITY_CLAMP.h>
#include <Hybrid_Methods/Forces/MPM_DRUCKER_PRAGER.h>
#include <Deformables/Constitutive_Models/ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.h>
#include <Hybrid_Methods/Forces/MPM_FORCE_HELPER.h>
#include <Hybrid_Methods/Iterators/GATHER_SCATTER.h>
#include <Hybrid_Methods/Iterators/PARTICLE_GRID_WEIGHTS_SPLINE.h>
#include <Hybrid_Methods/System/MPM_KRYLOV_SYSTEM.h>
#include <Hybrid_Methods/System/MPM_KRYLOV_VECTOR.h>
#include <Hybrid_Methods/System/MPM_OBJECTIVE.h>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef USE_OPENMP
#include <omp.h>
#endif
using namespace PhysBAM;

namespace{
struct READER
{
    FILE* f;
    explicit READER(const char* path):f(fopen(path,"rb")){if(!f){fprintf(stderr,"cannot open %s\n",path);exit(2);}}
    ~READER(){fclose(f);}
    template<class X> X get(){X x;if(fread(&x,sizeof(X),1,f)!=1){fprintf(stderr,"short read\n");exit(2);}return x;}
    template<class X> void get(X* p,size_t n){if(n && fread(p,sizeof(X),n,f)!=n){fprintf(stderr,"short read\n");exit(2);}}
};
struct WRITER
{
    FILE* f;
    explicit WRITER(const char* path):f(fopen(path,"wb")){if(!f){fprintf(stderr,"cannot open %s\n",path);exit(2);}}
    ~WRITER(){fclose(f);}
    template<class X> void put(const X& x){fwrite(&x,sizeof(X),1,f);}
    template<class X> void put(const X* p,size_t n){if(n) fwrite(p,sizeof(X),n,f);}
};

struct WALL {int type;double friction;double lo[3],hi[3];};
struct SCENE
{
    int dim,prec,model,nsteps,dump_every,nwalls,threads,plastic;
    int N[3];
    double lo[3],hi[3],gravity[3],dt,neo_threshold;
    double theta_c,theta_s,max_hardening,hardening_factor;  // clamped plasticity (plastic == 1)
    double dp[6];  // Drucker-Prager (plastic == 2): a0, a1, a3, a4, sigma_Y0, use_vc
    int reflect=0,full_reflect=0,side_type[6]={0,0,0,0,0,0};  // from the command line
    double reflect_friction=0;
    long long np;
    std::vector<WALL> walls;
    std::vector<double> X,V,F,B,mass,volume,mu,lambda;
    std::vector<unsigned char> valid;
};

SCENE Read_Scene(const char* path)
{
    READER r(path);
    char magic[8];r.get(magic,8);
    if(memcmp(magic,"MPMSCN1",7)){fprintf(stderr,"bad scene magic\n");exit(2);}
    SCENE s;
    s.dim=r.get<int>();s.prec=r.get<int>();s.model=r.get<int>();s.nsteps=r.get<int>();
    s.dump_every=r.get<int>();s.nwalls=r.get<int>();s.threads=r.get<int>();s.plastic=r.get<int>();
    r.get(s.N,3);r.get(s.lo,3);r.get(s.hi,3);r.get(s.gravity,3);
    s.dt=r.get<double>();s.neo_threshold=r.get<double>();
    s.theta_c=s.theta_s=s.max_hardening=s.hardening_factor=0;
    if(s.plastic==1){s.theta_c=r.get<double>();s.theta_s=r.get<double>();s.max_hardening=r.get<double>();
        s.hardening_factor=r.get<double>();}
    if(s.plastic==2) r.get(s.dp,6);
    s.np=r.get<long long>();
    for(int i=0;i<s.nwalls;i++){
        WALL w;w.type=r.get<int>();r.get<int>();w.friction=r.get<double>();
        r.get(w.lo,3);r.get(w.hi,3);s.walls.push_back(w);}
    size_t n=s.np,d=s.dim;
    s.X.resize(n*d);s.V.resize(n*d);s.F.resize(n*d*d);s.B.resize(n*d*d);
    s.mass.resize(n);s.volume.resize(n);s.mu.resize(n);s.lambda.resize(n);s.valid.resize(n);
    r.get(s.X.data(),n*d);r.get(s.V.data(),n*d);r.get(s.F.data(),n*d*d);r.get(s.B.data(),n*d*d);
    r.get(s.mass.data(),n);r.get(s.volume.data(),n);r.get(s.mu.data(),n);r.get(s.lambda.data(),n);
    r.get(s.valid.data(),n);
    return s;
}

template<class TV>
class HARNESS_EXAMPLE:public MPM_EXAMPLE<TV>
{
public:
    explicit HARNESS_EXAMPLE(const STREAM_TYPE st):MPM_EXAMPLE<TV>(st){}
    void Initialize() override {}
};

template<class TV>
struct RUN
{
    typedef typename TV::SCALAR T;
    enum {d=TV::m};
    typedef VECTOR<int,d> TV_INT;
    HARNESS_EXAMPLE<TV> ex;
    MPM_DRIVER<TV>* driver=0;
    MPM_DRUCKER_PRAGER<TV>* dp=0;
    const SCENE& s;

    RUN(const SCENE& s,STREAM_TYPE st):ex(st),s(s)
    {
        TV_INT N;TV lo,hi,g;
        for(int a=0;a<d;a++){N(a)=s.N[a];lo(a)=(T)s.lo[a];hi(a)=(T)s.hi[a];g(a)=(T)s.gravity[a];}
        ex.grid.Initialize(N,RANGE<TV>(lo,hi),true);
        ex.threads=s.threads;
        ex.gather_scatter.threads=s.threads;
        ex.Set_Weights(2);
        ex.use_symplectic_euler=true;
        ex.use_affine=true;
        ex.flip=0;
        MPM_PARTICLES<TV>& P=ex.particles;
        P.Store_B(true);
        P.Store_Lame(true);
        P.Add_Elements((int)s.np);
        for(int p=0;p<(int)s.np;p++){
            for(int a=0;a<d;a++){P.X(p)(a)=(T)s.X[p*d+a];P.V(p)(a)=(T)s.V[p*d+a];}
            for(int k=0;k<d*d;k++){P.F(p).x[k]=(T)s.F[p*d*d+k];P.B(p).x[k]=(T)s.B[p*d*d+k];}
            P.mass(p)=(T)s.mass[p];P.volume(p)=(T)s.volume[p];
            P.mu(p)=(T)s.mu[p];P.lambda(p)=(T)s.lambda[p];P.valid(p)=s.valid[p]!=0;}
        ISOTROPIC_CONSTITUTIVE_MODEL<T,d>* cm=0;
        if(s.model==0) cm=new COROTATED_FIXED<T,d>((T)1e4,(T).3);
        else if(s.model==2) cm=new ST_VENANT_KIRCHHOFF_HENCKY_STRAIN<T,d>((T)1e4,(T).3);
        else cm=new NEO_HOOKEAN<T,d>((T)1e4,(T).3,(T).05,(T)s.neo_threshold);
        MPM_FINITE_ELEMENTS<TV>* fe=new MPM_FINITE_ELEMENTS<TV>(ex.force_helper,*cm,ex.gather_scatter,0);
        ex.Add_Force(*fe);
        if(s.plastic==2){
            // STANDARD_TESTS_BASE::Add_Drucker_Prager (explicit plasticity), Fp stored
            P.Store_Fp(true);
            for(int p=0;p<(int)s.np;p++) P.Fp(p)=MATRIX<T,d>::Identity_Matrix();
            dp=new MPM_DRUCKER_PRAGER<TV>(P,&fe->gather_scatter,(T)s.dp[0],(T)s.dp[1],(T)s.dp[2],(T)s.dp[3]);
            dp->use_vc=s.dp[5]!=0;
            dp->Initialize_Particles(0,(T)s.dp[4]);
            ex.plasticity_models.Append(dp);}
        if(s.plastic==1){
            // STANDARD_TESTS_BASE::Add_Clamped_Plasticity (symplectic, explicit plasticity)
            P.Store_Fp(true);
            for(int p=0;p<(int)s.np;p++){
                P.Fp(p)=MATRIX<T,d>::Identity_Matrix();
                P.mu0(p)=P.mu(p);P.lambda0(p)=P.lambda(p);}
            MPM_PLASTICITY_CLAMP<TV>* pl=new MPM_PLASTICITY_CLAMP<TV>(P,&fe->gather_scatter,(T)s.theta_c,(T)s.theta_s,
                (T)s.max_hardening,(T)s.hardening_factor);
            ex.plasticity_models.Append(pl);}
        ex.Add_Force(*new MPM_GRAVITY<TV>(ex.force_helper,g,ex.gather_scatter,0));
        if(s.reflect){  // as -use_reflect [-use_reflect_friction] with Add_Walls (STANDARD_TESTS_BASE.cpp:487-511)
            ex.reflection_bc=true;
            ex.use_full_reflection=s.full_reflect!=0;
            ex.reflection_bc_friction=(T)s.reflect_friction;
            for(int i=0;i<2*d;i++) ex.side_bc_type(i)=(typename MPM_EXAMPLE<TV>::BC_TYPE)s.side_type[i];}
        for(const WALL& w:s.walls){
            TV a,b;for(int k=0;k<d;k++){a(k)=(T)w.lo[k];b(k)=(T)w.hi[k];}
            ex.Add_Collision_Object(RANGE<TV>(a,b),(typename MPM_EXAMPLE<TV>::COLLISION_TYPE)w.type,(T)w.friction);}

        driver=new MPM_DRIVER<TV>(ex);
        // Same set-up as MPM_DRIVER::Initialize, minus restart handling and frame output.
        auto dom=ex.grid.Domain_Indices(ex.ghost);
        ex.mass.Resize(dom);
        ex.velocity.Resize(dom);
        ex.velocity_save.Resize(dom);
        ex.velocity_friction_save.Resize(dom);
        driver->dv.u.Resize(dom);
        driver->rhs.u.Resize(dom);
        driver->objective.system.tmp.u.Resize(dom);
        P.Store_B(ex.use_affine);
        P.Store_S(ex.use_oldroyd);
        ex.Dp_inv.Resize(P.X.m);
        RANGE<TV_INT> range(ex.grid.Cell_Indices(ex.ghost));
        ex.location.Resize(range,no_init);
        for(RANGE_ITERATOR<d> it(range);it.Valid();it.Next())
            ex.location(it.index)=ex.grid.Center(it.index);
        driver->Update_Simulated_Particles();
    }
    ~RUN(){delete driver;}

    void Step()
    {
        ex.dt=(T)s.dt;
        driver->Advance_One_Time_Step();
        ex.time+=ex.dt;
    }

    void Dump(WRITER& w,int step)
    {
        w.put(step);
        MPM_PARTICLES<TV>& P=ex.particles;
        int n=P.X.m;
        std::vector<double> buf;
        buf.resize((size_t)n*d);
        for(int p=0;p<n;p++) for(int a=0;a<d;a++) buf[p*d+a]=P.X(p)(a);
        w.put(buf.data(),buf.size());
        for(int p=0;p<n;p++) for(int a=0;a<d;a++) buf[p*d+a]=P.V(p)(a);
        w.put(buf.data(),buf.size());
        buf.resize((size_t)n*d*d);
        for(int p=0;p<n;p++) for(int k=0;k<d*d;k++) buf[p*d*d+k]=P.F(p).x[k];
        w.put(buf.data(),buf.size());
        for(int p=0;p<n;p++) for(int k=0;k<d*d;k++) buf[p*d*d+k]=P.B(p).x[k];
        w.put(buf.data(),buf.size());
        std::vector<unsigned char> v(n);
        for(int p=0;p<n;p++) v[p]=P.valid(p);
        w.put(v.data(),v.size());
        int nn=ex.mass.array.m;
        buf.resize(nn);
        for(int i=0;i<nn;i++) buf[i]=ex.mass.array(i);
        w.put(buf.data(),buf.size());
        const ARRAY<TV,TV_INT>* arrs[3]={&ex.velocity,&ex.velocity_save,&ex.velocity_friction_save};
        buf.resize((size_t)nn*d);
        for(int k=0;k<3;k++){
            for(int i=0;i<nn;i++) for(int a=0;a<d;a++) buf[i*d+a]=arrs[k]->array(i)(a);
            w.put(buf.data(),buf.size());}
        if(s.plastic){  // plastic deformation and hardened Lame parameters
            buf.resize((size_t)n*d*d);
            for(int p=0;p<n;p++) for(int k=0;k<d*d;k++) buf[p*d*d+k]=P.Fp(p).x[k];
            w.put(buf.data(),buf.size());
            buf.resize(n);
            for(int p=0;p<n;p++) buf[p]=P.mu(p);
            w.put(buf.data(),buf.size());
            for(int p=0;p<n;p++) buf[p]=P.lambda(p);
            w.put(buf.data(),buf.size());}
        if(s.plastic==2){  // plastic_def, rho_F, sigma_Y, vc
            const ARRAY_VIEW<T>* a[4]={&dp->plastic_def,&dp->rho_F,&dp->sigma_Y,&dp->vc};
            for(int k=0;k<4;k++){
                for(int p=0;p<n;p++) buf[p]=(*a[k])(p);
                w.put(buf.data(),n);}}
    }
};

}
