//#####################################################################
// Reference harness: drives PhysBAM's own MPM_DRIVER (compiled from the
// sources under the reference checkout, never copied here) on a scene read
// from a binary file and dumps particle and grid state.  It is test
// infrastructure only: it runs in the authoring container to produce golden
// vectors and, as a binary under oracle/_ref/, to time the reference CPU step.
//
// Modes
//   ref_harness step  <scene.bin> <out.bin> [reflect <full> <friction> <side types...>]
//     (with "reflect": reflection_bc on, use_full_reflection = full,
//     reflection_bc_friction, side_bc_type of -x, +x, -y, +y[, -z, +z])
//   ref_harness time  <scene.bin> <threads> <warmup> <steps>
//   ref_harness svd   <prec> <dim> <in.bin> <out.bin>
//   ref_harness pstrain <prec> <dim> <model> <threshold> <in.bin> <out.bin>
//   ref_harness frames <scene.bin> <nframes> <frame_dt> <cfl> <min_dt> <max_dt> <out.bin>
//               [<strong_cfl> <sound_cfl> <cfl_F> <cfl_sound>]
//   ref_harness write <scene.bin> <dir>
//     (the scene's nsteps steps, then MPM_EXAMPLE::Write_Output_Files's
//     particle and restart files: <dir>/deformable_object_particles.gz, the
//     PARTICLES<TV>::Write stream, and <dir>/restart_data.gz, the time)
//   ref_harness read <prec> <dim> <dir> <out.bin>
//     (Read_From_File of <dir>/deformable_object_particles into MPM_PARTICLES
//     and of <dir>/restart_data: a dump of X, V, F, B, valid, mass, volume,
//     mu, lambda, Fp, mu0, lambda0 when stored, and the time)
//   ref_harness poisson <in.bin> <out.bin>
//     (STANDARD_TESTS_BASE::Seed_Particles_Poisson's sampling: RANDOM_NUMBERS
//     seeded once, POISSON_DISK with Set_Distance_By_Volume(dX.Product() /
//     particles_per_cell) of the MAC grid, Sample per shape in order; in.bin:
//     int prec, dim, seed, nshapes, cells[3]; double ppc, lo[3], hi[3]; per
//     shape int type (0 box, 1 sphere), double a[3], b[3]; out.bin: int count
//     per shape, then the points as doubles)
//   ref_harness limited <scene.bin> <strong_cfl> <sound_cfl> <cfl> <cfl_F> <cfl_sound> <min_dt> <out.bin>
//     (the scene's nsteps fixed-dt steps with MPM_DRIVER::Reduce_Dt: the dt
//     each step took and the final state, in the frames output layout)
// A scene with plasticity flag 1 adds MPM_PLASTICITY_CLAMP (Fp, mu0, lambda0
// stored), flag 2 MPM_DRUCKER_PRAGER; their dumps end with Fp, mu and lambda
// (and for flag 2 plastic_def, rho_F, sigma_Y, vc).  Model 2 is the Hencky
// strain model.
//   ref_harness weights <prec> <dim> <N> <lo> <hi> <in.bin> <out.bin>
//   ref_harness totals <scene.bin> <out.bin>
//     (the scene's nsteps steps with the reductions that -print_stats reports,
//     taken through the reference's own named-step callbacks: after
//     "update-weights" the particle totals and MPM_EXAMPLE::Potential_Energy,
//     after "p2g" and after "forces" the grid totals of example.velocity; then
//     the particle totals of the final state.  Records of 11 doubles: mass,
//     linear momentum (3), angular momentum (3; 2D in the first), kinetic
//     energy, potential energy, its MPM_FINITE_ELEMENTS and MPM_GRAVITY parts)
//
// Scene file layout (little endian) is written by tests/golden/scene_io.py.
//#####################################################################
#include "ref_scene.h"

namespace{
template<class TV>
int Run_Step(const SCENE& s,const char* out,STREAM_TYPE st)
{
#ifdef USE_OPENMP
    omp_set_num_threads(s.threads>0?s.threads:1);
#endif
    RUN<TV> run(s,st);
    WRITER w(out);
    w.put("MPMOUT1",8);
    int ndumps=s.dump_every>0?s.nsteps/s.dump_every:0;
    int nn=run.ex.mass.array.m;
    int np=(int)s.np;
    w.put(ndumps);w.put(s.dim);w.put(np);w.put(nn);
    for(int k=1;k<=s.nsteps;k++){
        run.Step();
        if(s.dump_every>0 && k%s.dump_every==0) run.Dump(w,k);}
    return 0;
}

// Totals of the reference's MPM_EXAMPLE (MPM_EXAMPLE.cpp:123-133, :209-315),
// recorded through Add_Callbacks around the named steps of Advance_One_Time_Step.
template<class TV>
int Run_Totals(const SCENE& s,const char* out,STREAM_TYPE st)
{
    typedef typename TV::SCALAR T;
    enum {d=TV::m};
#ifdef USE_OPENMP
    omp_set_num_threads(s.threads>0?s.threads:1);
#endif
    RUN<TV> run(s,st);
    auto& ex=run.ex;
    std::vector<double> rec;
    auto put_spin=[&](const typename TV::SPIN& a){for(int i=0;i<3;i++) rec.push_back(i<a.m?(double)a(i):0.0);};
    auto put_tv=[&](const TV& a){for(int i=0;i<3;i++) rec.push_back(i<d?(double)a(i):0.0);};
    auto particles=[&](){
        T m=0;
        for(int k=0;k<ex.simulated_particles.m;k++) m+=ex.particles.mass(ex.simulated_particles(k));
        rec.push_back(m);
        put_tv(ex.Total_Particle_Linear_Momentum());
        put_spin(ex.Total_Particle_Angular_Momentum());
        rec.push_back(ex.Total_Particle_Kinetic_Energy());
        ex.Capture_Stress();
        ex.Precompute_Forces(ex.time,ex.dt,false);
        rec.push_back(ex.Potential_Energy(ex.time));
        rec.push_back(ex.forces(0)->Potential_Energy(ex.time));
        rec.push_back(ex.forces(1)->Potential_Energy(ex.time));};
    auto grid=[&](){
        T m=0;
        for(int k=0;k<ex.valid_grid_indices.m;k++) m+=ex.mass.array(ex.valid_grid_indices(k));
        rec.push_back(m);
        put_tv(ex.Total_Grid_Linear_Momentum(ex.velocity));
        put_spin(ex.Total_Grid_Angular_Momentum(ex.dt,ex.velocity,0));
        rec.push_back(ex.Total_Grid_Kinetic_Energy(ex.velocity));
        for(int i=0;i<3;i++) rec.push_back(0);};
    ex.Add_Callbacks(false,"update-weights",particles);
    ex.Add_Callbacks(false,"p2g",grid);
    ex.Add_Callbacks(false,"forces",grid);
    for(int k=0;k<s.nsteps;k++) run.Step();
    run.driver->Update_Simulated_Particles();
    run.driver->Update_Particle_Weights();
    particles();
    WRITER w(out);
    int nrec=(int)(rec.size()/11);
    w.put(nrec);
    w.put(rec.data(),rec.size());
    return 0;
}

// Frames with the reference's adaptive dt: the substep loop of
// MPM_DRIVER::Simulate_To_Frame (MPM_DRIVER.cpp:173-197) around the
// reference's Compute_Dt and Advance_One_Time_Step, without callbacks and
// frame output.  Writes the substep dts and the final state.
template<class TV>
int Run_Frames(const SCENE& s,int nframes,double frame_dt,double cfl,double min_dt,double max_dt,const char* out,
    STREAM_TYPE st,const double* lim)
{
    typedef typename TV::SCALAR T;
#ifdef USE_OPENMP
    omp_set_num_threads(s.threads>0?s.threads:1);
#endif
    RUN<TV> run(s,st);
    auto& ex=run.ex;
    ex.frame_dt=(T)frame_dt;ex.cfl=(T)cfl;ex.min_dt=(T)min_dt;ex.max_dt=(T)max_dt;
    if(lim){ex.use_strong_cfl=lim[0]!=0;ex.use_sound_speed_cfl=lim[1]!=0;ex.cfl_F=(T)lim[2];ex.cfl_sound=(T)lim[3];}
    std::vector<double> dts;
    for(int f=0;f<nframes;f++){
        T time_at_frame=ex.time+ex.frame_dt;
        bool done=false;
        for(int substep=0;!done;substep++){
            ex.dt=run.driver->Compute_Dt();
            ex.dt=clamp(ex.dt,ex.min_dt,ex.max_dt);
            T next_time=ex.time+ex.dt;
            if(next_time+(T).0001*ex.dt>time_at_frame){next_time=time_at_frame;done=true;}
            else if(next_time+ex.dt>time_at_frame) next_time=(ex.time+time_at_frame)/2;
            ex.dt=next_time-ex.time;
            T original_dt=ex.dt;
            run.driver->Advance_One_Time_Step();
            if(ex.dt<original_dt){next_time=ex.time+ex.dt;done=false;}  // Reduce_Dt lowered it
            dts.push_back((double)ex.dt);
            ex.time=next_time;}}
    WRITER w(out);
    w.put("MPMFRM1",8);
    int nd=(int)dts.size(),np=(int)s.np,nn=run.ex.mass.array.m;
    w.put(nd);w.put(s.dim);w.put(np);w.put(nn);
    w.put(dts.data(),dts.size());
    run.Dump(w,nd);
    return 0;
}

// Steps, then the particle and restart files of Write_Output_Files
// (MPM_EXAMPLE.cpp:64-74; DEFORMABLE_BODY_COLLECTION.cpp:112).
template<class TV>
int Run_Write(const SCENE& s,const char* dir,STREAM_TYPE st)
{
#ifdef USE_OPENMP
    omp_set_num_threads(s.threads>0?s.threads:1);
#endif
    RUN<TV> run(s,st);
    for(int k=0;k<s.nsteps;k++) run.Step();
    Write_To_File(st,std::string(dir)+"/deformable_object_particles",run.ex.particles);
    Write_To_File(st,std::string(dir)+"/restart_data",run.ex.time);
    return 0;
}

// MPM_EXAMPLE::Read_Output_Files (MPM_EXAMPLE.cpp:95-102) into fresh particles.
template<class TV>
int Run_Read(const char* dir,const char* out)
{
    typedef typename TV::SCALAR T;
    enum {d=TV::m};
    MPM_PARTICLES<TV> P;
    T time=0;
    Read_From_File(std::string(dir)+"/deformable_object_particles",P);
    Read_From_File(std::string(dir)+"/restart_data",time);
    WRITER w(out);
    int n=P.X.m;
    w.put(n);
    std::vector<double> buf;
    auto vec=[&](const ARRAY_VIEW<TV>& a){buf.assign((size_t)n*d,0);
        for(int p=0;p<n;p++) for(int k=0;k<d;k++) buf[p*d+k]=a(p)(k);w.put(buf.data(),buf.size());};
    auto mat=[&](const ARRAY_VIEW<MATRIX<T,d> >& a){buf.assign((size_t)n*d*d,0);
        for(int p=0;p<n;p++) for(int k=0;k<d*d;k++) buf[p*d*d+k]=a(p).x[k];w.put(buf.data(),buf.size());};
    auto sca=[&](const ARRAY_VIEW<T>& a){buf.assign(n,0);for(int p=0;p<n;p++) buf[p]=a(p);w.put(buf.data(),buf.size());};
    vec(P.X);vec(P.V);mat(P.F);
    // the flags store_B / store_Fp stay false after a read: look the arrays up by name
    const ARRAY_VIEW<MATRIX<T,d> >* Bp=P.template Get_Array<MATRIX<T,d> >("B");
    if(Bp) mat(*Bp); else {buf.assign((size_t)n*d*d,0);w.put(buf.data(),buf.size());}
    std::vector<unsigned char> v(n);for(int p=0;p<n;p++) v[p]=P.valid(p);w.put(v.data(),v.size());
    sca(P.mass);sca(P.volume);
    const ARRAY_VIEW<T>* mu=P.template Get_Array<T>("mu");
    const ARRAY_VIEW<T>* la=P.template Get_Array<T>("lambda");
    if(mu) sca(*mu); else {buf.assign(n,0);w.put(buf.data(),n);}
    if(la) sca(*la); else {buf.assign(n,0);w.put(buf.data(),n);}
    const ARRAY_VIEW<MATRIX<T,d> >* Fpp=P.template Get_Array<MATRIX<T,d> >("Fp");
    int has_fp=Fpp?1:0;w.put(has_fp);
    if(has_fp){mat(*Fpp);sca(*P.template Get_Array<T>("mu0"));sca(*P.template Get_Array<T>("lambda0"));}
    double t=(double)time;w.put(t);
    return 0;
}

template<class TV>
int Run_Poisson(READER& r,int seed,int nshapes,const int* cells,double ppc,const double* lo,const double* hi,
    const char* out)
{
    typedef typename TV::SCALAR T;
    enum {d=TV::m};
    VECTOR<int,d> N;TV L,H;
    for(int a=0;a<d;a++){N(a)=cells[a];L(a)=(T)lo[a];H(a)=(T)hi[a];}
    GRID<TV> grid(N,RANGE<TV>(L,H),true);
    RANDOM_NUMBERS<T> random(seed);
    POISSON_DISK<TV> pd(1);
    pd.Set_Distance_By_Volume(grid.dX.Product()/(T)ppc);
    std::vector<int> counts;
    std::vector<double> pts;
    for(int k=0;k<nshapes;k++){
        int type=r.get<int>();
        double a[3],b[3];r.get(a,3);r.get(b,3);
        TV A,B;for(int c=0;c<d;c++){A(c)=(T)a[c];B(c)=(T)b[c];}
        ARRAY<TV> X;
        if(type==0){ANALYTIC_IMPLICIT_OBJECT<RANGE<TV> > io(RANGE<TV>(A,B));pd.Sample(random,io,X);}
        else{ANALYTIC_IMPLICIT_OBJECT<SPHERE<TV> > io(SPHERE<TV>(A,(T)b[0]));pd.Sample(random,io,X);}
        counts.push_back(X.m);
        for(int i=0;i<X.m;i++) for(int c=0;c<d;c++) pts.push_back((double)X(i)(c));}
    WRITER w(out);
    w.put(counts.data(),counts.size());
    w.put(pts.data(),pts.size());
    return 0;
}

// Fixed-dt steps with Reduce_Dt (use_strong_cfl, use_sound_speed_cfl).
template<class TV>
int Run_Limited(const SCENE& s,const double* lim,double cfl,double min_dt,const char* out,STREAM_TYPE st)
{
    typedef typename TV::SCALAR T;
#ifdef USE_OPENMP
    omp_set_num_threads(s.threads>0?s.threads:1);
#endif
    RUN<TV> run(s,st);
    auto& ex=run.ex;
    ex.use_strong_cfl=lim[0]!=0;ex.use_sound_speed_cfl=lim[1]!=0;ex.cfl_F=(T)lim[2];ex.cfl_sound=(T)lim[3];
    ex.cfl=(T)cfl;ex.min_dt=(T)min_dt;
    std::vector<double> dts;
    for(int k=0;k<s.nsteps;k++){run.Step();dts.push_back((double)ex.dt);}
    WRITER w(out);
    w.put("MPMFRM1",8);
    int nd=(int)dts.size(),np=(int)s.np,nn=run.ex.mass.array.m;
    w.put(nd);w.put(s.dim);w.put(np);w.put(nn);
    w.put(dts.data(),dts.size());
    run.Dump(w,nd);
    return 0;
}

template<class TV>
int Run_Time(const SCENE& s,int threads,int warmup,int steps,STREAM_TYPE st)
{
#ifdef USE_OPENMP
    omp_set_num_threads(threads);
#endif
    SCENE t=s;t.threads=threads;
    RUN<TV> run(t,st);
    for(int k=0;k<warmup;k++) run.Step();
    auto t0=std::chrono::steady_clock::now();
    for(int k=0;k<steps;k++) run.Step();
    auto t1=std::chrono::steady_clock::now();
    double sec=std::chrono::duration<double>(t1-t0).count();
    int active=0;
    for(int p=0;p<run.ex.particles.X.m;p++) active+=run.ex.particles.valid(p);
    printf("{\"particles\": %lld, \"active\": %d, \"threads\": %d, \"steps\": %d, \"seconds\": %.6f, \"particle_steps_per_s\": %.6e}\n",
        s.np,active,threads,steps,sec,(double)s.np*steps/sec);
    return 0;
}

template<class T,int d>
int Run_Svd(const char* in,const char* out)
{
    READER r(in);
    long long n=r.get<long long>();
    std::vector<double> A(n*d*d);
    r.get(A.data(),A.size());
    WRITER w(out);
    std::vector<double> U(n*d*d),S(n*d),V(n*d*d);
    for(long long i=0;i<n;i++){
        MATRIX<T,d> M,u,v;DIAGONAL_MATRIX<T,d> D;
        for(int k=0;k<d*d;k++) M.x[k]=(T)A[i*d*d+k];
        M.Singular_Value_Decomposition(u,D,v);
        for(int k=0;k<d*d;k++){U[i*d*d+k]=u.x[k];V[i*d*d+k]=v.x[k];}
        for(int k=0;k<d;k++) S[i*d+k]=D.x(k);}
    w.put(U.data(),U.size());w.put(S.data(),S.size());w.put(V.data(),V.size());
    return 0;
}

template<class T,int d>
int Run_Pstrain(int model,double threshold,const char* in,const char* out)
{
    READER r(in);
    long long n=r.get<long long>();
    std::vector<double> S(n*d),mu(n),la(n);
    r.get(S.data(),S.size());r.get(mu.data(),n);r.get(la.data(),n);
    ARRAY<T> amu(n),ala(n);
    for(long long i=0;i<n;i++){amu(i)=(T)mu[i];ala(i)=(T)la[i];}
    ISOTROPIC_CONSTITUTIVE_MODEL<T,d>* cm=0;
    if(model==0) cm=new COROTATED_FIXED<T,d>((T)1e4,(T).3);
    else cm=new NEO_HOOKEAN<T,d>((T)1e4,(T).3,(T).05,(T)threshold);
    ARRAY_VIEW<T> vmu(amu),vla(ala);
    cm->mu=&vmu;cm->lambda=&vla;
    std::vector<double> P(n*d);
    for(long long i=0;i<n;i++){
        DIAGONAL_MATRIX<T,d> D;for(int k=0;k<d;k++) D.x(k)=(T)S[i*d+k];
        DIAGONAL_MATRIX<T,d> R=cm->P_From_Strain(D,(int)i);
        for(int k=0;k<d;k++) P[i*d+k]=R.x(k);}
    cm->mu=0;cm->lambda=0;
    delete cm;
    WRITER w(out);
    w.put(P.data(),P.size());
    return 0;
}

template<class T,int d>
int Run_Weights(int N,double lo,double hi,const char* in,const char* out)
{
    typedef VECTOR<T,d> TV;
    READER r(in);
    long long n=r.get<long long>();
    std::vector<double> X(n*d);
    r.get(X.data(),X.size());
    GRID<TV> grid(VECTOR<int,d>()+N,RANGE<TV>(TV()+(T)lo,TV()+(T)hi),true);
    PARTICLE_GRID_WEIGHTS_SPLINE<TV,2> wts(grid,1);
    std::vector<int> base(n*d);
    std::vector<double> w(n*3*d),dw(n*3*d);
    for(long long i=0;i<n;i++){
        TV x;for(int a=0;a<d;a++) x(a)=(T)X[i*d+a];
        typename PARTICLE_GRID_WEIGHTS_SPLINE<TV,2>::PRECOMPUTE_DATA pd;
        wts.Compute_Precompute_Data(pd,x);
        for(int a=0;a<d;a++){
            base[i*d+a]=pd.base(a);
            for(int k=0;k<3;k++){w[(i*3+k)*d+a]=pd.w(k)(a);dw[(i*3+k)*d+a]=pd.dw(k)(a);}}}
    WRITER wr(out);
    wr.put(base.data(),base.size());wr.put(w.data(),w.size());wr.put(dw.data(),dw.size());
    return 0;
}
}

int main(int argc,char* argv[])
{
    LOG::Initialize_Logging(true,true,0,false);
    if(argc<2){fprintf(stderr,"usage: ref_harness step|time|svd|pstrain|weights ...\n");return 2;}
    std::string mode=argv[1];
    if(mode=="step" || mode=="time"){
        bool refl=mode=="step" && argc>4 && std::string(argv[4])=="reflect";
        if((mode=="step" && argc!=4 && !refl) || (mode=="time" && argc!=6)){fprintf(stderr,"bad arguments\n");return 2;}
        SCENE s=Read_Scene(argv[2]);
        if(refl){
            if(argc!=7+2*s.dim){fprintf(stderr,"bad reflect arguments\n");return 2;}
            s.reflect=1;s.full_reflect=atoi(argv[5]);s.reflect_friction=atof(argv[6]);
            for(int i=0;i<2*s.dim;i++) s.side_type[i]=atoi(argv[7+i]);}
        int rc=2;
        if(s.prec==8){
            STREAM_TYPE st((double()));
            if(s.dim==2) rc=mode=="step"?Run_Step<VECTOR<double,2> >(s,argv[3],st):Run_Time<VECTOR<double,2> >(s,atoi(argv[3]),atoi(argv[4]),atoi(argv[5]),st);
            else rc=mode=="step"?Run_Step<VECTOR<double,3> >(s,argv[3],st):Run_Time<VECTOR<double,3> >(s,atoi(argv[3]),atoi(argv[4]),atoi(argv[5]),st);}
        else{
            STREAM_TYPE st((float()));
            if(s.dim==2) rc=mode=="step"?Run_Step<VECTOR<float,2> >(s,argv[3],st):Run_Time<VECTOR<float,2> >(s,atoi(argv[3]),atoi(argv[4]),atoi(argv[5]),st);
            else rc=mode=="step"?Run_Step<VECTOR<float,3> >(s,argv[3],st):Run_Time<VECTOR<float,3> >(s,atoi(argv[3]),atoi(argv[4]),atoi(argv[5]),st);}
        LOG::Finish_Logging();
        return rc;}
    if(mode=="totals" && argc==4){
        SCENE s=Read_Scene(argv[2]);
        int rc;
        if(s.prec==8){
            STREAM_TYPE st((double()));
            rc=s.dim==2?Run_Totals<VECTOR<double,2> >(s,argv[3],st):Run_Totals<VECTOR<double,3> >(s,argv[3],st);}
        else{
            STREAM_TYPE st((float()));
            rc=s.dim==2?Run_Totals<VECTOR<float,2> >(s,argv[3],st):Run_Totals<VECTOR<float,3> >(s,argv[3],st);}
        LOG::Finish_Logging();
        return rc;}
    if(mode=="frames" && (argc==9 || argc==13)){
        SCENE s=Read_Scene(argv[2]);
        int nf=atoi(argv[3]);double fdt=atof(argv[4]),cfl=atof(argv[5]),mind=atof(argv[6]),maxd=atof(argv[7]);
        double limv[4];
        const double* lim=0;
        if(argc==13){for(int i=0;i<4;i++) limv[i]=atof(argv[9+i]);lim=limv;}
        int rc;
        if(s.prec==8){
            STREAM_TYPE st((double()));
            rc=s.dim==2?Run_Frames<VECTOR<double,2> >(s,nf,fdt,cfl,mind,maxd,argv[8],st,lim):
                Run_Frames<VECTOR<double,3> >(s,nf,fdt,cfl,mind,maxd,argv[8],st,lim);}
        else{
            STREAM_TYPE st((float()));
            rc=s.dim==2?Run_Frames<VECTOR<float,2> >(s,nf,fdt,cfl,mind,maxd,argv[8],st,lim):
                Run_Frames<VECTOR<float,3> >(s,nf,fdt,cfl,mind,maxd,argv[8],st,lim);}
        LOG::Finish_Logging();
        return rc;}
    if(mode=="write" && argc==4){
        SCENE s=Read_Scene(argv[2]);
        int rc;
        if(s.prec==8){
            STREAM_TYPE st((double()));
            rc=s.dim==2?Run_Write<VECTOR<double,2> >(s,argv[3],st):Run_Write<VECTOR<double,3> >(s,argv[3],st);}
        else{
            STREAM_TYPE st((float()));
            rc=s.dim==2?Run_Write<VECTOR<float,2> >(s,argv[3],st):Run_Write<VECTOR<float,3> >(s,argv[3],st);}
        LOG::Finish_Logging();
        return rc;}
    if(mode=="read" && argc==6){
        int prec=atoi(argv[2]),dim=atoi(argv[3]);
        int rc;
        if(prec==8) rc=dim==2?Run_Read<VECTOR<double,2> >(argv[4],argv[5]):Run_Read<VECTOR<double,3> >(argv[4],argv[5]);
        else rc=dim==2?Run_Read<VECTOR<float,2> >(argv[4],argv[5]):Run_Read<VECTOR<float,3> >(argv[4],argv[5]);
        LOG::Finish_Logging();
        return rc;}
    if(mode=="poisson" && argc==4){
        READER r(argv[2]);
        int prec=r.get<int>(),dim=r.get<int>(),seed=r.get<int>(),ns=r.get<int>();
        int cells[3];r.get(cells,3);
        double ppc=r.get<double>(),lo[3],hi[3];r.get(lo,3);r.get(hi,3);
        int rc;
        if(prec==8) rc=dim==2?Run_Poisson<VECTOR<double,2> >(r,seed,ns,cells,ppc,lo,hi,argv[3]):
            Run_Poisson<VECTOR<double,3> >(r,seed,ns,cells,ppc,lo,hi,argv[3]);
        else rc=dim==2?Run_Poisson<VECTOR<float,2> >(r,seed,ns,cells,ppc,lo,hi,argv[3]):
            Run_Poisson<VECTOR<float,3> >(r,seed,ns,cells,ppc,lo,hi,argv[3]);
        LOG::Finish_Logging();
        return rc;}
    if(mode=="limited" && argc==10){
        SCENE s=Read_Scene(argv[2]);
        double lim[4]={atof(argv[3]),atof(argv[4]),atof(argv[6]),atof(argv[7])};
        double cfl=atof(argv[5]),mind=atof(argv[8]);
        int rc;
        if(s.prec==8){
            STREAM_TYPE st((double()));
            rc=s.dim==2?Run_Limited<VECTOR<double,2> >(s,lim,cfl,mind,argv[9],st):
                Run_Limited<VECTOR<double,3> >(s,lim,cfl,mind,argv[9],st);}
        else{
            STREAM_TYPE st((float()));
            rc=s.dim==2?Run_Limited<VECTOR<float,2> >(s,lim,cfl,mind,argv[9],st):
                Run_Limited<VECTOR<float,3> >(s,lim,cfl,mind,argv[9],st);}
        LOG::Finish_Logging();
        return rc;}
    if(mode=="svd" && argc==6){
        int prec=atoi(argv[2]),dim=atoi(argv[3]);
        if(prec==8) return dim==2?Run_Svd<double,2>(argv[4],argv[5]):Run_Svd<double,3>(argv[4],argv[5]);
        return dim==2?Run_Svd<float,2>(argv[4],argv[5]):Run_Svd<float,3>(argv[4],argv[5]);}
    if(mode=="pstrain" && argc==8){
        int prec=atoi(argv[2]),dim=atoi(argv[3]),model=atoi(argv[4]);double thr=atof(argv[5]);
        if(prec==8) return dim==2?Run_Pstrain<double,2>(model,thr,argv[6],argv[7]):Run_Pstrain<double,3>(model,thr,argv[6],argv[7]);
        return dim==2?Run_Pstrain<float,2>(model,thr,argv[6],argv[7]):Run_Pstrain<float,3>(model,thr,argv[6],argv[7]);}
    if(mode=="weights" && argc==9){
        int prec=atoi(argv[2]),dim=atoi(argv[3]),N=atoi(argv[4]);double lo=atof(argv[5]),hi=atof(argv[6]);
        if(prec==8) return dim==2?Run_Weights<double,2>(N,lo,hi,argv[7],argv[8]):Run_Weights<double,3>(N,lo,hi,argv[7],argv[8]);
        return dim==2?Run_Weights<float,2>(N,lo,hi,argv[7],argv[8]):Run_Weights<float,3>(N,lo,hi,argv[7],argv[8]);}
    fprintf(stderr,"bad arguments\n");
    return 2;
}
