//#####################################################################
// mpm_oracle — CPU restatement of PhysBAM's explicit APIC MPM step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
// path: only the test suite, the smoke check and the CPU-baseline leg of the
// benchmark may load it.  The product library never links or calls it.
//
// It restates, in the same operation order as the reference (serial scatter,
// threads=1), the functions below.  Citations are relative to the PhysBAM
// checkout (Public_Library/...).
//   weights      Hybrid_Methods/Iterators/PARTICLE_GRID_WEIGHTS_SPLINE.cpp:33-56, :93-105, :133-140, :152-158
//   svd3         Core/Matrices/MATRIX_3X3.cpp:36-236   (Gast et al. 2016 implicit-shift QR)
//   svd2         Core/Matrices/MATRIX_2X2.h:313-322, SYMMETRIC_MATRIX_2X2.h:364-375
//   P_hat        Deformables/Constitutive_Models/COROTATED_FIXED.cpp:59-65,
//                NEO_HOOKEAN.cpp:44-81 (Clamp_To_Hyperbola + P_From_Strain)
//   stress       Hybrid_Methods/Forces/MPM_FINITE_ELEMENTS.cpp:47-67, :84-99
//   gravity      Hybrid_Methods/Forces/MPM_GRAVITY.cpp:62-69
//   P2G          Hybrid_Methods/Examples_And_Drivers/MPM_DRIVER.cpp:245-306
//   forces       MPM_DRIVER.cpp:959-1060 (symplectic branch)
//   collisions   Hybrid_Methods/System/MPM_OBJECTIVE.cpp:150-224; Collisions/MPM_COLLISION_OBJECT.cpp:27-75;
//                Core/Math_Tools/RANGE.cpp:38-63 (box phi and normal)
//   friction     MPM_DRIVER.cpp:1064-1094
//   G2P          MPM_DRIVER.cpp:450-515
//   invalidation MPM_DRIVER.cpp:1219-1239, simulated list :1142-1157
//   totals       Examples_And_Drivers/MPM_EXAMPLE.cpp:209-315 (linear and angular momentum, kinetic
//                energy of particles and grid), :123-133 potential energy with MPM_FINITE_ELEMENTS.cpp:72-80,
//                MPM_GRAVITY.cpp:47-55 and the models' Energy_Density; MPM_DRIVER.cpp:1159-1213 (-print_stats)
//   plasticity   Forces/MPM_PLASTICITY_CLAMP.cpp:31-45 (Compute with hardening),
//                Forces/MPM_DRUCKER_PRAGER.cpp:64-94 (Compute), :123-131 (Update_Hardening),
//                :40-58 (Initialize_Particles); Hencky stress ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:60-70
//                Forces/MPM_PLASTICITY_MODEL.cpp:34-48 (Update_Particles),
//                MPM_DRIVER.cpp:156, :920-926 (after G2P, over the step's simulated particles)
//   adaptive dt  MPM_DRIVER.cpp:1098-1104 (Compute_Dt), :1125-1138 (Grid_V_Upper_Bound),
//                :173-197 (substep loop of Simulate_To_Frame), Core/Math_Tools/clamp.h:13-16
//
// Each piece is pinned in tests/ against the reference harness outputs
// (golden vectors under tests/golden/) and against closed forms.
//#####################################################################
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle{

template<class T,int d> struct Vec
{
    T x[d];
    Vec(){for(int i=0;i<d;i++) x[i]=0;}
    T& operator[](int i){return x[i];}
    const T& operator[](int i) const{return x[i];}
};
template<class T,int d> Vec<T,d> operator+(const Vec<T,d>& a,const Vec<T,d>& b){Vec<T,d> r;for(int i=0;i<d;i++) r[i]=a[i]+b[i];return r;}
template<class T,int d> Vec<T,d> operator-(const Vec<T,d>& a,const Vec<T,d>& b){Vec<T,d> r;for(int i=0;i<d;i++) r[i]=a[i]-b[i];return r;}
template<class T,int d> Vec<T,d> operator*(T s,const Vec<T,d>& a){Vec<T,d> r;for(int i=0;i<d;i++) r[i]=s*a[i];return r;}
template<class T,int d> Vec<T,d> operator*(const Vec<T,d>& a,T s){Vec<T,d> r;for(int i=0;i<d;i++) r[i]=a[i]*s;return r;}
template<class T,int d> T dot(const Vec<T,d>& a,const Vec<T,d>& b){T s=a[0]*b[0];for(int i=1;i<d;i++) s+=a[i]*b[i];return s;}
template<class T,int d> T magnitude(const Vec<T,d>& a){return std::sqrt(dot(a,a));}

// Column-major storage x[i+d*j], as MATRIX<T,d> (Core/Matrices/MATRIX_3X3.h:115-119).
template<class T,int d> struct Mat
{
    T x[d*d];
    Mat(){for(int i=0;i<d*d;i++) x[i]=0;}
    static Mat Identity(){Mat m;for(int i=0;i<d;i++) m.x[i*(d+1)]=1;return m;}
    T& operator()(int i,int j){return x[i+d*j];}
    const T& operator()(int i,int j) const{return x[i+d*j];}
    Vec<T,d> Row(int i) const{Vec<T,d> r;for(int j=0;j<d;j++) r[j]=x[i+d*j];return r;}
    Vec<T,d> Col(int j) const{Vec<T,d> r;for(int i=0;i<d;i++) r[i]=x[i+d*j];return r;}
    void Set_Row(int i,const Vec<T,d>& v){for(int j=0;j<d;j++) x[i+d*j]=v[j];}
    void Set_Col(int j,const Vec<T,d>& v){for(int i=0;i<d;i++) x[i+d*j]=v[i];}
};
template<class T,int d> Mat<T,d> operator*(const Mat<T,d>& A,const Mat<T,d>& B)
{Mat<T,d> C;for(int j=0;j<d;j++) for(int i=0;i<d;i++){T s=A(i,0)*B(0,j);for(int k=1;k<d;k++) s+=A(i,k)*B(k,j);C(i,j)=s;}return C;}
template<class T,int d> Vec<T,d> operator*(const Mat<T,d>& A,const Vec<T,d>& v)
{Vec<T,d> r;for(int i=0;i<d;i++){T s=A(i,0)*v[0];for(int k=1;k<d;k++) s+=A(i,k)*v[k];r[i]=s;}return r;}
template<class T,int d> Mat<T,d> Times_Transpose(const Mat<T,d>& A,const Mat<T,d>& B)
{Mat<T,d> C;for(int j=0;j<d;j++) for(int i=0;i<d;i++){T s=A(i,0)*B(j,0);for(int k=1;k<d;k++) s+=A(i,k)*B(j,k);C(i,j)=s;}return C;}
template<class T,int d> Mat<T,d> operator*(T s,const Mat<T,d>& A){Mat<T,d> r;for(int i=0;i<d*d;i++) r.x[i]=s*A.x[i];return r;}
template<class T,int d> Mat<T,d> operator+(const Mat<T,d>& A,const Mat<T,d>& B){Mat<T,d> r;for(int i=0;i<d*d;i++) r.x[i]=A.x[i]+B.x[i];return r;}
template<class T,int d> Mat<T,d> Outer(const Vec<T,d>& u,const Vec<T,d>& v){Mat<T,d> r;for(int j=0;j<d;j++) for(int i=0;i<d;i++) r(i,j)=u[i]*v[j];return r;}
template<class T,int d> T Frobenius_Norm_Squared(const Mat<T,d>& A){T s=A.x[0]*A.x[0];for(int i=1;i<d*d;i++) s+=A.x[i]*A.x[i];return s;}

//#####################################################################
// 3x3 SVD: MATRIX_3X3.cpp:36-236.  U, V rotations; |s0|>=|s1|>=|s2|; only s2 may be negative.
//#####################################################################
template<class T> Vec<T,2> Givens(T a,T b)
{
    Vec<T,2> v;v[0]=a;v[1]=-b;
    T m=std::sqrt(a*a+b*b);
    if(m){T s=1/m;v[0]=a*s;v[1]=-b*s;}
    else{v[0]=1;v[1]=0;}
    return v;
}
template<class T> void Apply_L(Mat<T,3>& A,const Vec<T,2>& u,int i,int j)
{
    Vec<T,3> v=A.Row(i),w=A.Row(j);
    A.Set_Row(i,u[0]*v+u[1]*w);
    A.Set_Row(j,u[0]*w-u[1]*v);
}
template<class T> void Apply_R(Mat<T,3>& A,const Vec<T,2>& u,int i,int j)
{
    Vec<T,3> v=A.Col(i),w=A.Col(j);
    A.Set_Col(i,u[0]*v+u[1]*w);
    A.Set_Col(j,u[0]*w-u[1]*v);
}
template<class T> void Zero_L(Mat<T,3>& U,Mat<T,3>& A,int r0,int r1,int c)
{
    Vec<T,2> g=Givens(A(r0,c),A(r1,c));
    Apply_L(A,g,r1,r0);
    Apply_R(U,g,r1,r0);
    A(r1,c)=0;
}
template<class T> void Zero_R(Mat<T,3>& V,Mat<T,3>& A,int c0,int c1,int r)
{
    Vec<T,2> g=Givens(A(r,c0),A(r,c1));
    Apply_R(A,g,c1,c0);
    Apply_R(V,g,c1,c0);
    A(r,c1)=0;
}
template<class T> void Zero_Chasing(Mat<T,3>& U,Mat<T,3>& A,Mat<T,3>& V)
{
    Zero_L(U,A,0,1,0);
    Zero_R(V,A,1,2,0);
    Zero_L(U,A,1,2,1);
}
template<class T> void Swap_Col(Mat<T,3>& U,int i,int j)
{
    Vec<T,3> u=U.Col(i),v=U.Col(j);
    U.Set_Col(i,v);
    U.Set_Col(j,(T)-1*u);
}
template<class T> void Swap_Row(Mat<T,3>& U,int i,int j)
{
    Vec<T,3> u=U.Row(i),v=U.Row(j);
    U.Set_Row(i,v);
    U.Set_Row(j,(T)-1*u);
}
template<class T> void Swap(Mat<T,3>& U,Vec<T,3>& D,Mat<T,3>& V,int i,int j)
{
    Swap_Col(U,i,j);
    Swap_Col(V,i,j);
    T t=D[i];D[i]=D[j];D[j]=t;
}
template<class T> void Flip_Sign(Mat<T,3>& U,Vec<T,3>& D,int i,int j)
{
    D[i]=-D[i];D[j]=-D[j];
    U.Set_Col(i,(T)-1*U.Col(i));
    U.Set_Col(j,(T)-1*U.Col(j));
}
// argmax(a,b,c) of Core/Math_Tools/argmax.h: ties go to the earlier index.
template<class T> int Arg_Max3(T a,T b,T c){return a>=c?(a>=b?0:1):(b>=c?1:2);}
template<class T> void Sort_SV(Mat<T,3>& U,Vec<T,3>& D,Mat<T,3>& V)
{
    int i=Arg_Max3(std::abs(D[0]),std::abs(D[1]),std::abs(D[2]));
    if(i>0) Swap(U,D,V,0,i);
    if(std::abs(D[1])<std::abs(D[2])) Swap(U,D,V,1,2);
    if(D[0]<0) Flip_Sign(U,D,0,D[1]<0?1:2);
    else if(D[1]<0) Flip_Sign(U,D,1,2);
}

//#####################################################################
// 2x2 SVD: MATRIX_2X2.h:313-322 (indefinite polar + symmetric eigenproblem).
//#####################################################################
template<class T> void Svd2(const Mat<T,2>& A,Mat<T,2>& U,Vec<T,2>& sv,Mat<T,2>& V)
{
    const T* x=A.x;
    T x03=x[0]+x[3],c,s;
    if(x03==0){c=0;s=1;}
    else{T t=(x[1]-x[2])/x03;c=1/std::sqrt(1+t*t);s=t*c;}
    Mat<T,2> Q;Q.x[0]=c;Q.x[1]=s;Q.x[2]=-s;Q.x[3]=c;
    T s00=Q.x[0]*x[0]+Q.x[1]*x[1],s10=Q.x[0]*x[2]+Q.x[1]*x[3],s11=Q.x[2]*x[2]+Q.x[3]*x[3];
    // SYMMETRIC_MATRIX_2X2.h:364-375
    T a=(T).5*(s00+s11),b=(T).5*(s00-s11),cc=s10;
    T c2=cc*cc,m=std::sqrt(b*b+c2),k=s00*s11-c2;
    Vec<T,2> ev;
    if(a>=0){ev[0]=a+m;ev[1]=ev[0]?k/ev[0]:0;}
    else{ev[1]=a-m;ev[0]=ev[1]?k/ev[1]:0;}
    if(ev[1]>ev[0]){T t=ev[0];ev[0]=ev[1];ev[1]=t;}
    Vec<T,2> c0;
    if(b>=0){c0[0]=m+b;c0[1]=cc;}else{c0[0]=-cc;c0[1]=b-m;}
    T mag=std::sqrt(c0[0]*c0[0]+c0[1]*c0[1]);
    if(mag){T inv=1/mag;c0[0]*=inv;c0[1]*=inv;}else{c0[0]=1;c0[1]=0;}
    V.x[0]=c0[0];V.x[1]=c0[1];V.x[2]=-c0[1];V.x[3]=c0[0];
    sv=ev;
    if(sv[1]<0 && std::abs(sv[1])>=std::abs(sv[0])){
        T s0=-sv[1],s1=-sv[0];sv[0]=s0;sv[1]=s1;
        for(int i=0;i<4;i++) Q.x[i]=-Q.x[i];
        Mat<T,2> W;W.x[0]=V.x[2];W.x[1]=V.x[3];W.x[2]=-V.x[0];W.x[3]=-V.x[1];V=W;}
    U=Q*V;
}
template<class T> void Top_Block(Mat<T,3>& U,Mat<T,3>& B,Mat<T,3>& V,Vec<T,3>& D)
{
    Mat<T,2> a,u,v;Vec<T,2> dd;
    a.x[0]=B(0,0);a.x[1]=B(1,0);a.x[2]=B(0,1);a.x[3]=B(1,1);
    Svd2(a,u,dd,v);
    Vec<T,2> uc;uc[0]=u.x[0];uc[1]=u.x[1];
    Vec<T,2> vc;vc[0]=v.x[0];vc[1]=v.x[1];
    Apply_R(U,uc,0,1);
    Apply_R(V,vc,0,1);
    D[0]=dd[0];D[1]=dd[1];D[2]=B(2,2);
    Sort_SV(U,D,V);
}
template<class T> void Bot_Block(Mat<T,3>& U,Mat<T,3>& B,Mat<T,3>& V,Vec<T,3>& D)
{
    Mat<T,2> a,u,v;Vec<T,2> dd;
    a.x[0]=B(1,1);a.x[1]=B(2,1);a.x[2]=B(1,2);a.x[3]=B(2,2);
    Svd2(a,u,dd,v);
    Vec<T,2> uc;uc[0]=u.x[0];uc[1]=u.x[1];
    Vec<T,2> vc;vc[0]=v.x[0];vc[1]=v.x[1];
    Apply_R(U,uc,1,2);
    Apply_R(V,vc,1,2);
    D[0]=B(0,0);D[1]=dd[0];D[2]=dd[1];
    Sort_SV(U,D,V);
}
template<class T> bool Check_Tolerances(Mat<T,3>& B,Mat<T,3>& U,Vec<T,3>& D,Mat<T,3>& V,T tau)
{
    if(std::abs(B(1,2))<=tau){Top_Block(U,B,V,D);return true;}
    if(std::abs(B(0,1))<=tau){Bot_Block(U,B,V,D);return true;}
    if(std::abs(B(1,1))<=tau){
        Vec<T,2> g=Givens(B(1,2),B(2,2));
        Vec<T,2> o;o[0]=-g[1];o[1]=g[0];
        Apply_L(B,o,2,1);
        Apply_R(U,o,2,1);
        Top_Block(U,B,V,D);
        return true;}
    if(std::abs(B(2,2))<=tau){
        Zero_R(V,B,1,2,1);
        Zero_R(V,B,0,2,0);
        Top_Block(U,B,V,D);
        return true;}
    if(std::abs(B(0,0))<=tau){
        Zero_L(U,B,1,0,1);
        Zero_L(U,B,2,0,2);
        Bot_Block(U,B,V,D);
        return true;}
    return false;
}
template<class T> bool Lazy_Sort(Mat<T,3>& U,Mat<T,3>& B,Mat<T,3>& V,int i,int j)
{
    T flip_ratio=100;
    if(std::abs(B(i,i))*flip_ratio>=std::abs(B(j,j))) return false;
    Swap_Row(B,i,j);
    Swap_Col(U,i,j);
    Swap_Col(B,i,j);
    Swap_Col(V,i,j);
    return true;
}
template<class T> void Bidiagonalize(Mat<T,3>& U,Mat<T,3>& B,Mat<T,3>& V)
{
    Zero_L(U,B,1,2,0);
    Zero_Chasing(U,B,V);
}
template<class T> int Svd3(const Mat<T,3>& F,Mat<T,3>& U,Vec<T,3>& D,Mat<T,3>& V)
{
    Mat<T,3> B=F;
    U=V=Mat<T,3>::Identity();
    Bidiagonalize(U,B,V);
    T nu=16*std::numeric_limits<T>::epsilon();
    T half_norm=(T).5*std::sqrt(Frobenius_Norm_Squared(B));
    T tau=nu*(half_norm<(T)1?(T)1:half_norm);
    for(int it=0;it<100000;it++){
        if(Check_Tolerances(B,U,D,V,tau)) return it;
        bool s01=Lazy_Sort(U,B,V,0,1);
        bool s02=Lazy_Sort(U,B,V,0,2);
        bool s12=Lazy_Sort(U,B,V,1,2);
        if(s01|s02|s12){
            Bidiagonalize(U,B,V);
            if(Check_Tolerances(B,U,D,V,tau)) return it;}
        T al0=B(0,0),al1=B(1,1),al2=B(2,2);
        T be0=B(0,1),be1=B(1,2);
        T ga0=al0*be0,ga1=al1*be1;
        T a0=al1*al1+be0*be0;
        T a1=al2*al2+be1*be1;
        T b0=ga1;
        T dd=(T).5*(a0-a1);
        T sgn=dd>=0?(T)1:(T)-1;
        T mu=a1-b0*b0/(dd+sgn*std::sqrt(dd*dd+b0*b0));
        Vec<T,2> g=Givens(al0*al0-mu,ga0);
        Apply_R(B,g,1,0);
        Apply_R(V,g,1,0);
        Zero_Chasing(U,B,V);}
    return -1;
}
template<class T> int Svd(const Mat<T,3>& F,Mat<T,3>& U,Vec<T,3>& D,Mat<T,3>& V){return Svd3(F,U,D,V);}
template<class T> int Svd(const Mat<T,2>& F,Mat<T,2>& U,Vec<T,2>& D,Mat<T,2>& V){Svd2(F,U,D,V);return 0;}

//#####################################################################
// Constitutive models in singular-value space.
//#####################################################################
// DIAGONAL_MATRIX::Cofactor_Matrix (DIAGONAL_MATRIX.h:204-219): 2D (y,x); 3D (y*z, z*x, y*x).
template<class T> Vec<T,2> Cofactor(const Vec<T,2>& s){Vec<T,2> r;r[0]=s[1];r[1]=s[0];return r;}
template<class T> Vec<T,3> Cofactor(const Vec<T,3>& s){Vec<T,3> r;T yz=s[1]*s[2];r[0]=yz;r[1]=s[2]*s[0];r[2]=s[1]*s[0];return r;}
template<class T,int d> T Product(const Vec<T,d>& s){T p=s[0];for(int i=1;i<d;i++) p*=s[i];return p;}

// COROTATED_FIXED.cpp:59-65
template<class T,int d> Vec<T,d> P_Corotated(const Vec<T,d>& s,T mu,T la)
{
    Vec<T,d> cof=Cofactor(s);
    T J=Product(s);
    T a=2*mu,b=la*(J-1);
    Vec<T,d> r;
    for(int i=0;i<d;i++) r[i]=(s[i]-1)*a+cof[i]*b;
    return r;
}
// NEO_HOOKEAN.cpp:44-68 Clamp_To_Hyperbola
template<class T> Vec<T,2> Clamp_To_Hyperbola(const Vec<T,2>& F,T thr,T root)
{
    Vec<T,2> r;
    if(F[0]*F[0]>thr){r[0]=F[0];r[1]=thr/F[0];}
    else{r[0]=root;r[1]=root;}
    return r;
}
template<class T> Vec<T,3> Clamp_To_Hyperbola(const Vec<T,3>& F,T thr,T root)
{
    Vec<T,3> r;
    if(F[0]*F[1]*F[1]>thr){r[0]=F[0];r[1]=F[1];r[2]=thr/(F[0]*F[1]);}
    else if(F[0]*F[0]*F[0]>thr){T c=std::sqrt(thr/F[0]);r[0]=F[0];r[1]=c;r[2]=c;}
    else{r[0]=root;r[1]=root;r[2]=root;}
    return r;
}
// pow<1,d>(t) of the constructor: the d-th root of the threshold.
template<class T> T Dth_Root(T t,int d){return d==2?std::sqrt(t):std::cbrt(t);}
// NEO_HOOKEAN.cpp:73-81
// ST_VENANT_KIRCHHOFF_HENCKY_STRAIN::P_From_Strain (ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:60-70):
// ((2 mu) ln s + lambda tr(ln s)) (1/s), zero if the last singular value is not positive.
template<class T,int d> Vec<T,d> P_Hencky(const Vec<T,d>& s,T mu,T la)
{
    Vec<T,d> P;
    if(s[d-1]<=0){for(int i=0;i<d;i++) P[i]=0;return P;}
    Vec<T,d> l;for(int i=0;i<d;i++) l[i]=std::log(s[i]);
    T tr=l[0];for(int i=1;i<d;i++) tr+=l[i];
    T c=la*tr,m2=(T)2*mu;
    for(int i=0;i<d;i++) P[i]=(m2*l[i]+c)*((T)1/s[i]);
    return P;
}

template<class T,int d> Vec<T,d> P_Neo_Hookean(const Vec<T,d>& s,T mu,T la,T thr)
{
    T J=Product(s);
    Vec<T,d> r;
    if(J>=thr){
        T c=mu-la*std::log(J);
        for(int i=0;i<d;i++) r[i]=s[i]*mu-(1/s[i])*c;
        return r;}
    Vec<T,d> Fc=Clamp_To_Hyperbola(s,thr,Dth_Root(thr,d)),dF,Fi;
    for(int i=0;i<d;i++){Fi[i]=1/Fc[i];dF[i]=s[i]-Fc[i];}
    T c=mu-la*std::log(thr);
    T ip=Fi[0]*dF[0];for(int i=1;i<d;i++) ip+=Fi[i]*dF[i];
    T li=la*ip;
    for(int i=0;i<d;i++){
        T a=s[i]*mu;
        T b=(Fi[i]*Fi[i]*dF[i]-Fi[i])*c;
        r[i]=a+b+Fi[i]*li;}
    return r;
}

// The parts of Isotropic_Stress_Derivative that Sound_Speed reads: the
// diagonal of H and the B (spin) entries, for the three models.
//   fixed corotated COROTATED_FIXED.cpp:127-140 (2D), :144-167 (3D), panic_threshold 1e-6 (:22)
//   neo-Hookean     NEO_HOOKEAN.cpp:143-156 (2D), :161-180 (3D)
//   Hencky          ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:85-102 (2D), :106-137 (3D), failure_threshold 0 (:28)
// log1p_x_over_x / diff_log_over_diff: Core/Math_Tools/Robust_Functions.h:33-39.
template<class T> T Log1p_X_Over_X(T x){if(std::fabs(x)<1e-18) return 1;return std::log1p(x)/x;}
template<class T> T Diff_Log_Over_Diff(T x,T y){return Log1p_X_Over_X((y-x)/x)/x;}
template<class T> T Panic(T d){const T pt=(T)1e-6;if(std::fabs(d)<pt) d=d<0?-pt:pt;return d;}
template<class T> void Stress_Derivative_HB(int model,const Vec<T,2>& F,T mu,T la,T thr,Vec<T,2>& H,T* B)
{
    if(model==0){
        T mu2=2*mu,d01=Panic(F[0]+F[1]);
        H[0]=mu2+la*(F[1]*F[1]);H[1]=mu2+la*(F[0]*F[0]);
        B[0]=mu2*(1-1/d01);}
    else if(model==1){
        Vec<T,2> Fc=(F[0]*F[1]>=thr)?F:Clamp_To_Hyperbola(F,thr,Dth_Root(thr,2));
        Vec<T,2> Fi;for(int a=0;a<2;a++) Fi[a]=1/Fc[a];
        T mlj=mu+la*std::log(Fi[0]*Fi[1]);
        for(int a=0;a<2;a++) H[a]=mu+(la+mlj)*(Fi[a]*Fi[a]);
        B[0]=mu;}
    else{
        if(F[1]<=0){H[0]=H[1]=B[0]=0;return;}
        T s0=std::max(F[0],(T)0),s1=std::max(F[1],(T)0),lp2m=la+2*mu,two_mu=2*mu;
        T l0=std::log(s0),l1=std::log(s1),sum_log=l0+l1;
        T x00=s0*s0,x11=s1*s1,x10=s1*s0;
        T m01=(two_mu/s0)*(Diff_Log_Over_Diff(s0,s1)-l1/s1)-la*sum_log/x10;
        T p01=((la*s0+lp2m*s1)*l0+(la*s1+lp2m*s0)*l1)/(x10*(s0+s1));
        H[0]=(lp2m*(1-l0)-la*l1)/x00;H[1]=(lp2m*(1-l1)-la*l0)/x11;
        B[0]=(T)0.5*(m01+p01);}
}
template<class T> void Stress_Derivative_HB(int model,const Vec<T,3>& F,T mu,T la,T thr,Vec<T,3>& H,T* B)
{
    if(model==0){
        T mu2=2*mu;
        Vec<T,3> cof=Cofactor(F);
        T d01=Panic(F[0]+F[1]),d02=Panic(F[0]+F[2]),d12=Panic(F[1]+F[2]);
        for(int a=0;a<3;a++) H[a]=mu2+la*(cof[a]*cof[a]);
        B[2]=mu2*(1-1/d01);B[1]=mu2*(1-1/d02);B[0]=mu2*(1-1/d12);}
    else if(model==1){
        Vec<T,3> Fc=(Product(F)>=thr)?F:Clamp_To_Hyperbola(F,thr,Dth_Root(thr,3));
        Vec<T,3> Fi;for(int a=0;a<3;a++) Fi[a]=1/Fc[a];
        T mlj=mu+la*std::log(Product(Fi));
        for(int a=0;a<3;a++) H[a]=mu+(la+mlj)*(Fi[a]*Fi[a]);
        B[0]=B[1]=B[2]=mu;}
    else{
        if(F[2]<=0){H[0]=H[1]=H[2]=B[0]=B[1]=B[2]=0;return;}
        T s[3],l[3];
        for(int a=0;a<3;a++){s[a]=std::max(F[a],(T)0);l[a]=std::log(s[a]);}
        T lp2m=la+2*mu,two_mu=2*mu,sum_log=l[0]+l[1]+l[2];
        T row[3];for(int i=0;i<3;i++) row[i]=s[i]*l[0]+s[i]*l[1]+s[i]*l[2];
        for(int a=0;a<3;a++) H[a]=(lp2m-two_mu*l[a]-la*sum_log)/(s[a]*s[a]);
        // B(2) from the pair (0,1), B(1) from (0,2), B(0) from (1,2)
        const int pi[3]={1,0,0},pj[3]={2,2,1};
        for(int k=0;k<3;k++){
            int i=pi[k],j=pj[k];
            T xij=s[j]*s[i];
            T mij=(two_mu/s[i])*(Diff_Log_Over_Diff(s[i],s[j])-l[j]/s[j])-la*sum_log/xij;
            T pij=(two_mu*(s[j]*l[i]+s[i]*l[j])+la*(row[i]+row[j]))/(xij*(s[i]+s[j]));
            B[k]=(T)0.5*(mij+pij);}}
}
// ISOTROPIC_CONSTITUTIVE_MODEL::Sound_Speed (ISOTROPIC_CONSTITUTIVE_MODEL.cpp:151-162):
// sqrt(max(H_ii F_i^2, B_a F_{(a+1)%m}^2, B_a F_{(a+2)%m}^2) / density), m the spin size.
template<class T,int d> T Sound_Speed(int model,const Vec<T,d>& F,T mu,T la,T thr,T density)
{
    Vec<T,d> H;T B[3];
    Stress_Derivative_HB(model,F,mu,la,thr,H,B);
    Vec<T,d> F2;for(int a=0;a<d;a++) F2[a]=F[a]*F[a];
    T cm=H[0]*F2[0];for(int a=1;a<d;a++) cm=std::max(cm,H[a]*F2[a]);
    const int m=d==2?1:3;
    for(int a=0;a<m;a++) for(int i=0;i<2;i++) cm=std::max(cm,B[a]*F2[(a+i+1)%m]);
    return std::sqrt(cm/density);
}

// Energy_Density of the singular values F (what MPM_FINITE_ELEMENTS::Potential_Energy
// sums, MPM_FINITE_ELEMENTS.cpp:72-80):
//   fixed corotated COROTATED_FIXED.cpp:179-185: mu tr((F-1)^2) + lambda/2 (J-1)^2
//   neo-Hookean     NEO_HOOKEAN.cpp:191-199:     mu (tr(F F^T)/2 - d/2 - ln J) + lambda/2 ln^2 J, 1e16 for J <= 0
//   Hencky          ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:49-59: mu |ln F|^2 + lambda/2 tr(ln F)^2, FLT_MAX if F_d <= 0
template<class T,int d> T Energy_Density(int model,const Vec<T,d>& F,T mu,T la)
{
    if(model==0){
        T s=0;for(int i=0;i<d;i++){T f=F[i]-1;s+=f*f;}
        T j=Product(F)-1;
        return mu*s+(T).5*la*(j*j);}
    if(model==1){
        T I1=0;for(int i=0;i<d;i++) I1+=F[i]*F[i];
        T J=Product(F);
        if(J<=0) return (T)1e16;
        T lj=std::log(J);
        return mu*((T).5*(I1-d)-lj)+(T).5*la*(lj*lj);}
    if(F[d-1]<=0) return std::numeric_limits<float>::max();
    Vec<T,d> l;for(int i=0;i<d;i++) l[i]=std::log(F[i]);
    T tr=0,fn=0;for(int i=0;i<d;i++){tr+=l[i];fn+=l[i]*l[i];}
    return mu*fn+(T)0.5*la*(tr*tr);
}
// Core/Math_Tools/Robust_Arithmetic.h:22-26 (FLT_MAX in T for both precisions).
template<class T> T Robust_Divide(T a,T b)
{
    const T fm=(T)std::numeric_limits<float>::max();
    T aa=std::fabs(a),ab=std::fabs(b);
    if(ab==fm) return T();
    if(ab>1 || aa<fm*ab) return a/b;
    return (a>0)==(b>=0)?fm:-fm;
}
// Enforce_Limit_Max (MPM_DRIVER.cpp:549-559): lower s so that
// |t (a + t b)| <= bound for 0 <= t <= s.
template<class T> void Enforce_Limit_Max(T& s,T bound,T a,T b)
{
    if(a<0){a=-a;b=-b;}
    if(a*a+4*b*bound>=0){
        if(s*(a+s*b)>bound) s=(2*bound)/(a+std::sqrt(a*a+4*b*bound));}
    else{
        if(s*(a+s*b)<-bound) s=-(a+std::sqrt(a*a-4*b*bound))/(2*b);}
}

//#####################################################################
// Scene and state
//#####################################################################
struct Wall{int type;double friction;double lo[3],hi[3];};   // type: 0 stick, 1 slip, 2 separate

template<class T,int d> struct Scene
{
    int N[3];
    Vec<T,d> lo,hi,dX,one_over_dX,gravity;
    int model;T neo_threshold;T dt;
    std::vector<Wall> walls;
    // Reduce_Dt settings (MPM_EXAMPLE.h:95-104): example.cfl, cfl_F,
    // cfl_sound, min_dt, use_strong_cfl, use_sound_speed_cfl
    int strong_cfl=0,sound_cfl=0;
    T cfl=1,cfl_F=(T).5,cfl_sound=(T).9,min_dt=0;
    // reflection boundary conditions (MPM_EXAMPLE.h:106-109, :134):
    // reflection_bc, use_full_reflection, reflection_bc_friction and
    // side_bc_type per side -x, +x, -y, +y, -z, +z (0 free, 1 slip, 2 noslip, 3 separate)
    int reflect=0,full_reflect=0;
    T reflect_friction=0;
    int side_type[6]={0,0,0,0,0,0};
    int gn[3];        // nodes per axis including 3 ghost layers each side
    long long nn;     // total nodes
};

template<class T,int d> struct Box{Vec<T,d> lo,hi;};

// RANGE<TV>::Signed_Distance, RANGE.cpp:38-45
template<class T,int d> T Box_Phi(const Box<T,d>& b,const Vec<T,d>& X)
{
    Vec<T,d> len=b.hi-b.lo,c=(b.lo+b.hi)*(T).5,phi;
    bool all_le=true;
    for(int i=0;i<d;i++){phi[i]=std::abs(X[i]-c[i])-(T).5*len[i];if(!(phi[i]<=0)) all_le=false;}
    if(!all_le){Vec<T,d> p;for(int i=0;i<d;i++) p[i]=phi[i]>0?phi[i]:0;return magnitude(p);}
    T m=phi[0];for(int i=1;i<d;i++) if(phi[i]>m) m=phi[i];
    return m;
}
// RANGE<TV>::Normal, RANGE.cpp:50-63
template<class T,int d> Vec<T,d> Box_Normal(const Box<T,d>& b,const Vec<T,d>& X)
{
    bool inside=true;
    for(int i=0;i<d;i++) if(!(X[i]>=b.lo[i] && X[i]<=b.hi[i])) inside=false;
    Vec<T,d> pmax=X-b.hi,pmin=b.lo-X,n;
    if(inside){
        Vec<T,d> m;for(int i=0;i<d;i++) m[i]=pmin[i]>pmax[i]?pmin[i]:pmax[i];
        int axis=0;
        if(d==3) axis=Arg_Max3(m[0],m[1],m[2]);
        else axis=m[0]>=m[1]?0:1;
        n[axis]=pmax[axis]>pmin[axis]?(T)1:(T)-1;
        return n;}
    for(int i=0;i<d;i++){
        T phi=pmin[i]>pmax[i]?pmin[i]:pmax[i];
        n[i]=phi>0?(pmax[i]>pmin[i]?phi:-phi):0;}
    T m=magnitude(n);
    if(m) return n*(1/m);
    Vec<T,d> e;e[0]=1;return e;
}

template<class T,int d> struct Weights
{
    int base[d];
    T w[3][d],dw[3][d];
};

// PARTICLE_GRID_WEIGHTS_SPLINE.cpp:133-140 (Compute_Precompute_Data) with :93-105 (degree 2)
template<class T,int d> void Compute_Weights(const Scene<T,d>& s,const Vec<T,d>& X,Weights<T,d>& pd)
{
    Vec<T,d> Y=X-s.dX;  // X-(T).5*degree*DX with degree 2
    for(int a=0;a<d;a++) pd.base[a]=(int)std::floor((Y[a]-s.lo[a])*s.one_over_dX[a]);
    for(int a=0;a<d;a++){
        T center=s.lo[a]+((T)pd.base[a]+(T).5)*s.dX[a];
        T x=(X[a]-center)*s.one_over_dX[a];
        T q=x-(T)1.5;
        pd.w[0][a]=(T).5*(q*q);
        pd.dw[0][a]=q*s.one_over_dX[a];
        x-=(T)1;
        pd.w[1][a]=-x*x+(T).75;
        pd.dw[1][a]=((T)(-2)*x)*s.one_over_dX[a];
        x-=(T)1;
        T r=x+(T)1.5;
        pd.w[2][a]=(T).5*(r*r);
        pd.dw[2][a]=r*s.one_over_dX[a];}
}

// One stencil entry (PARTICLE_GRID_WEIGHTS_SPLINE.cpp:33-56); k enumerates {0,1,2}^d with the last axis fastest.
template<class T,int d> void Stencil(const Weights<T,d>& pd,int k,int idx[d],T& w,Vec<T,d>& g)
{
    int o[d];
    for(int a=d-1;a>=0;a--){o[a]=k%3;k/=3;}
    T aw[d];
    for(int a=0;a<d;a++){idx[a]=pd.base[a]+o[a];aw[a]=pd.w[o[a]][a];}
    w=aw[0];for(int a=1;a<d;a++) w*=aw[a];
    for(int i=0;i<d;i++) g[i]=1;
    for(int j=0;j<d;j++){
        T dj=pd.dw[o[j]][j];
        for(int i=0;i<d;i++) g[i]*=(i==j)?dj:aw[j];}
}

template<class T,int d> long long Flat(const Scene<T,d>& s,const int idx[d])
{
    long long f=0;
    for(int a=0;a<d;a++) f=f*s.gn[a]+(idx[a]+3);
    return f;
}
template<class T,int d> Vec<T,d> Node_Center(const Scene<T,d>& s,long long f)
{
    int idx[d];
    for(int a=d-1;a>=0;a--){idx[a]=(int)(f%s.gn[a])-3;f/=s.gn[a];}
    Vec<T,d> c;
    for(int a=0;a<d;a++) c[a]=s.lo[a]+((T)idx[a]+(T).5)*s.dX[a];
    return c;
}

template<class T,int d> struct Particles
{
    long long n;
    T* X;T* V;T* F;T* B;
    const T* mass;const T* volume;const T* mu;const T* lambda;
    unsigned char* valid;
    Vec<T,d> x(long long p) const{Vec<T,d> r;for(int a=0;a<d;a++) r[a]=X[p*d+a];return r;}
    Vec<T,d> v(long long p) const{Vec<T,d> r;for(int a=0;a<d;a++) r[a]=V[p*d+a];return r;}
    Mat<T,d> f(long long p) const{Mat<T,d> r;for(int k=0;k<d*d;k++) r.x[k]=F[p*d*d+k];return r;}
    Mat<T,d> b(long long p) const{Mat<T,d> r;for(int k=0;k<d*d;k++) r.x[k]=B[p*d*d+k];return r;}
};

template<class T,int d> struct Grid
{
    std::vector<T> mass;
    std::vector<Vec<T,d> > velocity,velocity_save,velocity_friction_save,v0,dv,tmp2,tmp0,tmp1,v1;
    std::vector<long long> valid;
};

struct Collision{int object;long long p;};

// Per-particle stress A_p = PFT (MPM_FINITE_ELEMENTS.cpp:59-65); volume applied by the caller.
template<class T,int d> Mat<T,d> Stress_PFT(const Scene<T,d>& s,const Mat<T,d>& F,const Mat<T,d>& Fn,T mu,T la)
{
    Mat<T,d> U,V;Vec<T,d> sig;
    Svd(F,U,sig,V);
    Mat<T,d> FV=Fn*V;
    Vec<T,d> P=s.model==0?P_Corotated(sig,mu,la):(s.model==2?P_Hencky(sig,mu,la):P_Neo_Hookean(sig,mu,la,s.neo_threshold));
    Mat<T,d> DT;  // DIAGONAL_MATRIX::Times_Transpose: r(i,j)=P(i)*FV(j,i)
    for(int i=0;i<d;i++) for(int j=0;j<d;j++) DT(i,j)=P[i]*FV(j,i);
    return U*DT;
}

//#####################################################################
// Reflection boundary conditions (MPM_DRIVER.cpp:310-446).  Each visits the
// ghost nodes of the domain side by side (-x, +x, -y, +y, -z, +z), each side
// over its full slab, corners included (RANGE_ITERATOR with ghost |
// duplicate_corners, Core/Math_Tools/RANGE_ITERATOR.cpp:125-150), pairing a
// ghost node with its mirror image across the side: index 2 r - i - 1 along
// the side's axis, r = 0 or N.
//#####################################################################
enum{BC_FREE=0,BC_SLIP=1,BC_NOSLIP=2,BC_SEP=3};
template<class T,int d,class FN> void For_Ghost_Pairs(const Scene<T,d>& s,int side,FN fn)
{
    const int axis=side/2,hi=side%2,N=s.N[axis];
    int lo[d],up[d];
    for(int a=0;a<d;a++){lo[a]=-3;up[a]=s.N[a]+3;}
    lo[axis]=hi?N:-3;up[axis]=hi?N+3:0;
    int idx[d];for(int a=0;a<d;a++) idx[a]=lo[a];
    while(true){
        int img[d];for(int a=0;a<d;a++) img[a]=idx[a];
        img[axis]=2*(hi?N:0)-idx[axis]-1;
        fn(Flat(s,img),Flat(s,idx),axis,hi);
        int a=d-1;
        while(a>=0 && ++idx[a]==up[a]){idx[a]=lo[a];a--;}
        if(a<0) break;}
}
// Reflect_Boundary_Mass_Momentum (:310-342) without a boundary velocity; the
// same map on a force field is Reflect_Boundary_Force (:422-446).
template<class T,int d> void Reflect_Pair_Vectors(int type,int axis,Vec<T,d>& vi,Vec<T,d>& vo)
{
    Vec<T,d> id=vi-vo,od=(T)-1*id,in=vi+vo,on=in;
    if(type==BC_SLIP){in[axis]=id[axis];on[axis]=od[axis];}
    if(type==BC_NOSLIP){vi=id;vo=od;}
    else{vi=in;vo=on;}
}
template<class T,int d> void Reflect_Boundary_Mass_Momentum(const Scene<T,d>& s,Grid<T,d>& G)
{
    for(int side=0;side<2*d;side++)
        For_Ghost_Pairs(s,side,[&](long long i,long long o,int axis,int hi){
            T mt=G.mass[i]+G.mass[o];
            Reflect_Pair_Vectors(s.side_type[side],axis,G.velocity[i],G.velocity[o]);
            G.mass[o]=G.mass[i]=mt;});
}
template<class T,int d> void Reflect_Boundary_Force(const Scene<T,d>& s,std::vector<Vec<T,d> >& f)
{
    for(int side=0;side<2*d;side++)
        For_Ghost_Pairs(s,side,[&](long long i,long long o,int axis,int hi){
            Reflect_Pair_Vectors(s.side_type[side],axis,f[i],f[o]);});
}
// Reflect_Boundary_Velocity (:395-420): the ghost node takes the mirrored
// velocity of its image (normal part negated for slip, all of it for noslip).
template<class T,int d> void Reflect_Boundary_Velocity(const Scene<T,d>& s,const Grid<T,d>& G,std::vector<Vec<T,d> >& u)
{
    for(int side=0;side<2*d;side++)
        For_Ghost_Pairs(s,side,[&](long long i,long long o,int axis,int hi){
            if(!G.mass[o]) return;
            Vec<T,d> vi=u[i],od=(T)-1*vi,on=vi;
            int type=s.side_type[side];
            if(type==BC_SLIP) on[axis]=od[axis];
            if(type==BC_NOSLIP) u[o]=od;
            else u[o]=on;});
}
// VECTOR::Normalize: divides by the magnitude (the x axis if zero) and returns it.
template<class T,int d> T Normalize(Vec<T,d>& v)
{
    T m=magnitude(v);
    if(m){T r=1/m;for(int a=0;a<d;a++) v[a]*=r;}
    else{v=Vec<T,d>();v[0]=1;}
    return m;
}
// Reflect_Helper (:343-369) without a boundary velocity.
template<class T,int d> void Reflect_Helper(T m_i,T m_o,Vec<T,d>& v_i,Vec<T,d>& v_o,int type,T mu,int a,int s)
{
    T m_t=m_i+m_o,a_i=m_i/m_t,a_o=1-a_i;
    Vec<T,d> v_in=a_i*v_i+a_o*v_o,v_on=v_in;
    if(type==BC_FREE){v_i=v_in;v_o=v_on;return;}
    Vec<T,d> bc;
    Vec<T,d> v_id=a_i*v_i+a_o*((T)2*bc-v_o),v_od=a_o*v_o+a_i*((T)2*bc-v_i);
    if(type==BC_NOSLIP){v_i=v_id;v_o=v_od;return;}
    if(type==BC_SLIP){v_in[a]=v_id[a];v_on[a]=v_od[a];v_i=v_in;v_o=v_on;return;}
    T d_v=2*a_o*a_i*(2*bc[a]-v_i[a]-v_o[a]);
    T d_v_n=s?-d_v:d_v;
    if(d_v_n<0) return;
    Vec<T,d> v_id_n,v_od_n;
    v_id_n[a]=v_id[a];
    v_od_n[a]=v_od[a];
    v_in[a]=0;
    T d_v_t_mag=Normalize(v_in);
    T r=d_v_t_mag-mu*d_v_n;
    if(r<0){v_i=v_id_n;v_o=v_od_n;return;}
    v_i=v_id_n+r*v_in;v_o=v_od_n+r*v_in;
}
// Reflect_Boundary_Friction (:371-393): velocity pairs where both nodes hold mass.
template<class T,int d> void Reflect_Boundary_Friction(const Scene<T,d>& s,Grid<T,d>& G)
{
    for(int side=0;side<2*d;side++)
        For_Ghost_Pairs(s,side,[&](long long i,long long o,int axis,int hi){
            T m_i=G.mass[i],m_o=G.mass[o];
            if(!m_i || !m_o) return;
            Reflect_Helper(m_i,m_o,G.velocity[i],G.velocity[o],s.side_type[side],s.reflect_friction,axis,hi);});
}

// MPM_DRIVER::Grid_To_Particle_Limit_Dt (MPM_DRIVER.cpp:572-620) without the
// midpoint rule: s = min over the simulated particles of the largest t <= 100
// with |t (dt grad v_old + t dt grad(v_new - v_old))| <= cfl_F per entry and
// |t (dt v_old + t dt (v_new - v_old))| <= cfl per axis (gathered with the
// step's weights); returns dt s.
template<class T,int d> T Grid_To_Particle_Limit_Dt(const Scene<T,d>& s,const Particles<T,d>& P,const Grid<T,d>& G,
    const std::vector<Weights<T,d> >& W,const std::vector<long long>& sim)
{
    const T dt=s.dt;
    const int K=d==3?27:9;
    T smin=100;
    for(long long p:sim){
        Vec<T,d> V_pic,V_pic_s;
        Mat<T,d> grad,grad_s;
        for(int k=0;k<K;k++){
            int idx[d];T w;Vec<T,d> g;
            Stencil(W[p],k,idx,w,g);
            long long f=Flat(s,idx);
            Vec<T,d> V_new=G.velocity[f],V_old=G.velocity_save[f];
            V_pic=V_pic+w*V_old;
            V_pic_s=V_pic_s+w*(V_new-V_old);
            Vec<T,d> V_grid_s=(T)1*(V_new-V_old);
            grad=grad+Outer(V_old,g);
            grad_s=grad_s+Outer(V_grid_s,g);}
        T sp=100;
        Mat<T,d> a=dt*grad,b=dt*grad_s;
        for(int i=0;i<d;i++) for(int j=0;j<d;j++) Enforce_Limit_Max(sp,s.cfl_F,a(i,j),b(i,j));
        Vec<T,d> xa=dt*V_pic,xb=dt*V_pic_s;
        for(int i=0;i<d;i++) Enforce_Limit_Max(sp,s.cfl,xa[i],xb[i]);
        smin=std::min(sp,smin);}
    return s.dt*smin;
}

// MPM_DRIVER::Reduce_Dt (MPM_DRIVER.cpp:519-545) with Limit_Dt_Sound_Speed
// (:624-637) and Compute_Max_Sound_Speed (:893-918): the sound speed of each
// simulated particle from the singular values of its F at the start of the
// step (MPM_FINITE_ELEMENTS::Precompute), density m / V0.  When the new dt is
// below 0.99 dt, the grid velocities on valid nodes are scaled toward
// velocity_save so the step's changes shrink with it.  Returns the dt G2P uses.
template<class T,int d> T Reduce_Dt(const Scene<T,d>& s,const Particles<T,d>& P,Grid<T,d>& G,
    const std::vector<Weights<T,d> >& W,const std::vector<long long>& sim)
{
    T dt=s.dt;
    if(s.strong_cfl){
        T new_dt=Grid_To_Particle_Limit_Dt(s,P,G,W,sim);
        if(new_dt<dt) dt=new_dt;}
    T dts=s.dt;
    if(s.sound_cfl){
        T max_speed=0;
        for(long long p:sim){
            Mat<T,d> U,V;Vec<T,d> sig;
            Svd(P.f(p),U,sig,V);
            T density=P.mass[p]/P.volume[p];
            max_speed=std::max(max_speed,Sound_Speed(s.model,sig,P.mu[p],P.lambda[p],s.neo_threshold,density));}
        T dxmin=s.dX[0];for(int a=1;a<d;a++) dxmin=std::min(dxmin,s.dX[a]);
        T new_dt=Robust_Divide(dxmin,max_speed)*s.cfl_sound;
        dts=std::min(dts,new_dt);}
    dt=std::min(dt,dts);
    dt=std::max(dt,s.min_dt);
    if(dt>=s.dt*(T)0.99) return s.dt;
    T sc=dt/s.dt;
    for(long long j:G.valid){
        Vec<T,d> vs=G.velocity_save[j];
        G.velocity[j]=(G.velocity[j]-vs)*sc+vs;
        G.velocity_friction_save[j]=(G.velocity_friction_save[j]-vs)*sc+vs;}
    return dt;
}

// One Advance_One_Time_Step; returns the dt it took (s.dt unless Reduce_Dt lowered it).
template<class T,int d> T Step(const Scene<T,d>& s,Particles<T,d>& P,Grid<T,d>& G,T time)
{
    const T dt=s.dt;
    // Update_Simulated_Particles (MPM_DRIVER.cpp:1142-1157)
    std::vector<long long> sim;
    for(long long p=0;p<P.n;p++) if(P.valid[p]) sim.push_back(p);
    std::vector<Weights<T,d> > W(P.n);
    for(long long p=0;p<P.n;p++) Compute_Weights(s,P.x(p),W[p]);
    T dpinv[d];
    for(int a=0;a<d;a++) dpinv[a]=((T)4*s.one_over_dX[a])*s.one_over_dX[a];
    const int K=d==3?27:9;

    // Particle_To_Grid (MPM_DRIVER.cpp:245-306)
    long long nn=s.nn;
    G.mass.assign(nn,0);
    G.velocity.assign(nn,Vec<T,d>());
    G.velocity_save.assign(nn,Vec<T,d>());
    G.velocity_friction_save.assign(nn,Vec<T,d>());
    if((long long)G.dv.size()!=nn){G.dv.assign(nn,Vec<T,d>());G.tmp0=G.tmp1=G.v1=G.dv;}
    for(long long p:sim){
        Vec<T,d> Xp=P.x(p),Vp=P.v(p);Mat<T,d> Bp=P.b(p);T mp=P.mass[p];
        for(int k=0;k<K;k++){
            int idx[d];T w;Vec<T,d> g;
            Stencil(W[p],k,idx,w,g);
            long long f=Flat(s,idx);
            G.mass[f]+=w*mp;
            Vec<T,d> Vv=w*Vp;
            Vec<T,d> Y=Node_Center(s,f)-Xp;
            Vec<T,d> DY;  // SYMMETRIC_MATRIX * VECTOR with zero off-diagonals
            for(int i=0;i<d;i++){T acc=(i==0?dpinv[0]:(T)0)*Y[0];for(int j=1;j<d;j++) acc+=(i==j?dpinv[i]:(T)0)*Y[j];DY[i]=acc;}
            Vv=Vv+(Bp*DY)*w;
            G.velocity[f]=G.velocity[f]+mp*Vv;}}
    if(s.reflect && !s.full_reflect) Reflect_Boundary_Mass_Momentum(s,G);
    G.valid.clear();
    for(long long i=0;i<nn;i++){
        if(G.mass[i]!=0){
            G.valid.push_back(i);
            Vec<T,d> v=G.velocity[i]*((T)1/G.mass[i]);
            G.velocity[i]=v;G.velocity_save[i]=v;}
        else G.velocity[i]=Vec<T,d>();}

    // Apply_Forces, symplectic branch (MPM_DRIVER.cpp:959-1060)
    G.v0=G.velocity;
    std::vector<Mat<T,d> > PFT(P.n);
    for(long long p:sim){
        Mat<T,d> F=P.f(p);
        PFT[p]=Stress_PFT(s,F,F,P.mu[p],P.lambda[p]);}
    G.tmp2.assign(nn,Vec<T,d>());
    for(long long p:sim){  // MPM_FINITE_ELEMENTS::Add_Forces
        Mat<T,d> A=P.volume[p]*PFT[p];
        for(int k=0;k<K;k++){
            int idx[d];T w;Vec<T,d> g;
            Stencil(W[p],k,idx,w,g);
            long long f=Flat(s,idx);
            G.tmp2[f]=G.tmp2[f]-A*g;}}
    for(long long p:sim){  // MPM_GRAVITY::Add_Forces
        for(int k=0;k<K;k++){
            int idx[d];T w;Vec<T,d> g;
            Stencil(W[p],k,idx,w,g);
            long long f=Flat(s,idx);
            G.tmp2[f]=G.tmp2[f]+(w*P.mass[p])*s.gravity;}}
    if(s.reflect && !s.full_reflect) Reflect_Boundary_Force(s,G.tmp2);
    for(long long i:G.valid) G.dv[i]=(dt/G.mass[i])*G.tmp2[i];
    for(long long i:G.valid) G.tmp1[i]=G.dv[i];

    // Adjust_For_Collision / Test_Add_Collision (MPM_OBJECTIVE.cpp:150-224)
    std::vector<Collision> collisions;
    std::vector<Vec<T,d> > coll_n;
    std::vector<long long> stuck;
    std::vector<Box<T,d> > boxes(s.walls.size());
    for(size_t o=0;o<s.walls.size();o++) for(int a=0;a<d;a++){boxes[o].lo[a]=(T)s.walls[o].lo[a];boxes[o].hi[a]=(T)s.walls[o].hi[a];}
    T t1=time+dt;(void)t1;
    for(long long i:G.valid){
        for(size_t o=0;o<s.walls.size();o++){
            Vec<T,d> Vn=G.v0[i]+G.dv[i];
            Vec<T,d> X0=Node_Center(s,i);
            Vec<T,d> X=X0+dt*Vn;
            T phi0=Box_Phi(boxes[o],X0);
            T phi=Box_Phi(boxes[o],X)-(phi0<0?phi0:(T)0);
            int type=s.walls[o].type;
            if(type==0 && phi0>0) continue;
            if(type==1 && phi0>0) continue;
            if(type==2 && phi>(T)1e-15) continue;
            if(type!=0 && phi>std::numeric_limits<float>::max()) continue;
            Vec<T,d> n=Box_Normal(boxes[o],X);
            collisions.push_back({(int)o,i});
            coll_n.push_back(n);
            X=X-phi*n;
            Vn=(X-X0)*((T)1/dt);
            if(type==0){Vn=Vec<T,d>();stuck.push_back(i);}
            G.dv[i]=(Vn-G.v0[i])*((T)1/(T)1);}}
    for(long long i:G.valid) G.tmp0[i]=G.dv[i];
    if(s.reflect && s.full_reflect){  // MPM_DRIVER.cpp:999-1019
        for(long long j:G.valid) G.velocity[j]=G.dv[j]+G.v0[j];
        Reflect_Boundary_Friction(s,G);
        for(long long j:G.valid) G.dv[j]=G.velocity[j]-G.v0[j];}

    // Apply_Friction (MPM_DRIVER.cpp:1064-1094)
    for(long long j:G.valid) G.velocity_friction_save[j]=G.dv[j];
    if(!s.walls.empty()){
        for(long long j:G.valid) G.v1[j]=(T)1*G.v0[j]+G.dv[j];
        for(long long j:stuck) G.v1[j]=Vec<T,d>();
        for(size_t c=0;c<collisions.size();c++){
            long long p=collisions[c].p;
            const Vec<T,d>& n=coll_n[c];
            Vec<T,d>& v=G.v1[p];
            T normal_force=dot(n,G.tmp0[p]-G.tmp1[p]);
            Vec<T,d> t=v-dot(v,n)*n;
            T t_mag=magnitude(t);
            if(t_mag){T inv=1/t_mag;t=t*inv;}else{t=Vec<T,d>();t[0]=1;}
            T k=(T)s.walls[collisions[c].object].friction*normal_force;
            if(t_mag<=k) v=dot(v,n)*n;
            else v=v-k*t;
            G.velocity_friction_save[p]=v-G.v0[p];}}
    for(long long j:G.valid){
        G.velocity[j]=G.dv[j]+G.v0[j];
        G.velocity_friction_save[j]=G.velocity_friction_save[j]+G.v0[j];}
    for(long long j:stuck) G.velocity_friction_save[j]=Vec<T,d>();
    if(s.reflect && !s.full_reflect){  // MPM_DRIVER.cpp:1057-1059
        Reflect_Boundary_Velocity(s,G,G.velocity);
        Reflect_Boundary_Velocity(s,G,G.velocity_friction_save);}

    const T dtg=(s.strong_cfl || s.sound_cfl)?Reduce_Dt(s,P,G,W,sim):dt;

    // Grid_To_Particle (MPM_DRIVER.cpp:450-515) + Reflect_Or_Invalidate_Particle (:1219-1239)
    Vec<T,d> A=s.lo,Bd=s.hi;
    Vec<T,d> eps=(std::numeric_limits<T>::epsilon()*(T)4)*(Bd-A);
    A=A+eps;Bd=Bd-eps;
    for(long long p:sim){
        Vec<T,d> V_pic,V_weight_old,V_pic_fric;
        Mat<T,d> Bh,grad;
        for(int k=0;k<K;k++){
            int idx[d];T w;Vec<T,d> g;
            Stencil(W[p],k,idx,w,g);
            long long f=Flat(s,idx);
            Vec<T,d> V_new=G.velocity[f],V_old=G.velocity_save[f],V_fric=G.velocity_friction_save[f];
            V_pic=V_pic+w*V_new;
            V_weight_old=V_weight_old+w*V_old;
            V_pic_fric=V_pic_fric+w*V_fric;
            grad=grad+Outer(V_new,g);
            Vec<T,d> Z=Node_Center(s,f);
            Vec<T,d> xi_new=Z+dtg*V_new;
            Bh=Bh+(w/2)*(Outer(V_fric,Z+xi_new)+Outer(Z-xi_new,V_fric));}
        Mat<T,d> Am=dtg*grad;
        for(int i=0;i<d;i++) Am(i,i)+=1;
        Mat<T,d> Fn=Am*P.f(p);
        Vec<T,d> Xp=P.x(p);
        Vec<T,d> xp_new=Xp+dtg*V_pic;
        Mat<T,d> corr=(T).5*(Outer(V_pic_fric,Xp+xp_new)+Outer(Xp-xp_new,V_pic_fric));
        for(int q=0;q<d*d;q++) Bh.x[q]-=corr.x[q];
        Vec<T,d> V_flip=(P.v(p)+V_pic_fric)-V_weight_old;
        Vec<T,d> Vnew=V_flip*(T)0+V_pic_fric*(T)1;
        for(int q=0;q<d*d;q++){P.F[p*d*d+q]=Fn.x[q];P.B[p*d*d+q]=Bh.x[q];}
        for(int a=0;a<d;a++){P.X[p*d+a]=xp_new[a];P.V[p*d+a]=Vnew[a];}
        for(int a=0;a<d;a++){  // sides with a boundary condition reflect the particle
            const bool rl=s.side_type[2*a]!=BC_FREE,rh=s.side_type[2*a+1]!=BC_FREE;
            T& x=P.X[p*d+a];
            T& v=P.V[p*d+a];
            if(x<A[a]){if(rl){x=2*A[a]-x;v=-v;}else P.valid[p]=0;}
            else if(x>Bd[a]){if(rh){x=2*Bd[a]-x;v=-v;}else P.valid[p]=0;}}}
    return dtg;
}

template<class T,int d> Scene<T,d> Make_Scene(const int* N,const double* lo,const double* hi,const double* g,int model,double thr,double dt,
    int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi)
{
    Scene<T,d> s;
    s.nn=1;
    for(int a=0;a<3;a++) s.N[a]=a<d?N[a]:1;
    for(int a=0;a<d;a++){
        s.lo[a]=(T)lo[a];s.hi[a]=(T)hi[a];s.gravity[a]=(T)g[a];
        s.dX[a]=(s.hi[a]-s.lo[a])/(T)N[a];
        s.one_over_dX[a]=(T)1/s.dX[a];
        s.gn[a]=N[a]+6;s.nn*=s.gn[a];}
    s.model=model;s.neo_threshold=(T)thr;s.dt=(T)dt;
    for(int i=0;i<nwalls;i++){
        Wall w;w.type=wtype[i];w.friction=wfric[i];
        for(int a=0;a<3;a++){w.lo[a]=wlo[3*i+a];w.hi[a]=whi[3*i+a];}
        s.walls.push_back(w);}
    return s;
}

//#####################################################################
// Conservation totals and energies, as -print_stats reports them
// (MPM_DRIVER.cpp:1159-1213).  Output layout (11 values): mass, linear
// momentum (d), angular momentum (SPIN: 1 value in 2D, 3 in 3D) at [4..6],
// kinetic energy [7], potential energy [8] = elastic [9] + gravity [10]
// (particles only).  Serial sums in the reference's order (threads = 1).
//#####################################################################
// TV::Cross_Product (VECTOR_2D.h:331-332, VECTOR_3D.h:321-322)
template<class T> void Cross(const Vec<T,2>& a,const Vec<T,2>& b,T* c){c[0]=a[0]*b[1]-a[1]*b[0];}
template<class T> void Cross(const Vec<T,3>& a,const Vec<T,3>& b,T* c)
{c[0]=a[1]*b[2]-a[2]*b[1];c[1]=a[2]*b[0]-a[0]*b[2];c[2]=a[0]*b[1]-a[1]*b[0];}
// MATRIX::Contract_Permutation_Tensor (MATRIX_2X2.h:339-340, MATRIX_3X3.h:392-393)
template<class T> void Contract_Permutation(const Mat<T,2>& B,T* c){c[0]=B.x[2]-B.x[1];}
template<class T> void Contract_Permutation(const Mat<T,3>& B,T* c){c[0]=B.x[7]-B.x[5];c[1]=B.x[2]-B.x[6];c[2]=B.x[3]-B.x[1];}

// Total_Particle_Linear_Momentum, Total_Particle_Angular_Momentum (with the
// -m B:eps term of store_B), Total_Particle_Kinetic_Energy (with the
// m/2 (B D^-1):B term, D^-1 = 4/dx^2 of the quadratic spline)
// (MPM_EXAMPLE.cpp:209-225, :248-263, :302-315) and Potential_Energy
// (MPM_EXAMPLE.cpp:123-133: MPM_FINITE_ELEMENTS then MPM_GRAVITY,
// MPM_FINITE_ELEMENTS.cpp:72-80 over the singular values of F, MPM_GRAVITY.cpp:47-55)
// over the valid particles.
template<class T,int d> void Particle_Totals(const Scene<T,d>& s,const Particles<T,d>& P,T* out)
{
    const int m=d==2?1:3;
    for(int i=0;i<11;i++) out[i]=0;
    T dpinv[d];
    for(int a=0;a<d;a++) dpinv[a]=((T)4*s.one_over_dX[a])*s.one_over_dX[a];
    T pe_fe=0,pe_g=0;
    for(long long p=0;p<P.n;p++){
        if(!P.valid[p]) continue;
        const T mp=P.mass[p];
        Vec<T,d> X=P.x(p),V=P.v(p);
        Mat<T,d> B=P.b(p);
        out[0]+=mp;
        for(int a=0;a<d;a++) out[1+a]+=mp*V[a];
        T c[3],e[3];
        Cross(X,V,c);Contract_Permutation(B,e);
        for(int a=0;a<m;a++){out[4+a]+=mp*c[a];out[4+a]-=mp*e[a];}
        T ke=mp/2*dot(V,V);
        // (B*D).Double_Contract(B): column by column, rows inside (MATRIX_BASE.h:446-449)
        T bb=0;
        for(int j=0;j<d;j++) for(int i=0;i<d;i++) bb+=(B(i,j)*dpinv[j])*B(i,j);
        ke+=mp/2*bb;
        out[7]+=ke;
        Mat<T,d> U,W;Vec<T,d> sig;
        Svd(P.f(p),U,sig,W);
        pe_fe+=P.volume[p]*Energy_Density(s.model,sig,P.mu[p],P.lambda[p]);
        pe_g-=mp*dot(X,s.gravity);}
    out[9]=pe_fe;out[10]=pe_g;
    T pe=0;pe+=pe_fe;pe+=pe_g;
    out[8]=pe;
}
// Total_Grid_Linear_Momentum, Total_Grid_Angular_Momentum (no midpoint),
// Total_Grid_Kinetic_Energy of the node velocities u over valid_grid_indices
// (MPM_EXAMPLE.cpp:226-241, :268-284, :289-297), node locations the cell
// centres of the MAC grid (MPM_DRIVER.cpp:115-126).
template<class T,int d> void Grid_Totals(const Scene<T,d>& s,const Grid<T,d>& G,const std::vector<Vec<T,d> >& u,T* out)
{
    const int m=d==2?1:3;
    for(int i=0;i<11;i++) out[i]=0;
    for(long long i:G.valid){
        const T mi=G.mass[i];
        out[0]+=mi;
        for(int a=0;a<d;a++) out[1+a]+=mi*u[i][a];
        T c[3];
        Cross(Node_Center(s,i),u[i],c);
        for(int a=0;a<m;a++) out[4+a]+=mi*c[a];
        out[7]+=(T).5*mi*dot(u[i],u[i]);}
}

template<class T,int d> int Run(const int* N,const double* lo,const double* hi,const double* g,int model,double thr,double dt,
    int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi,
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps,
    T* gmass,T* gvel,T* gvsave,T* gvfric)
{
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi);
    Particles<T,d> P{np,X,V,F,B,mass,volume,mu,la,valid};
    Grid<T,d> G;
    T time=0;
    for(int k=0;k<nsteps;k++) time+=Step(s,P,G,time);
    if(nsteps>0){
        for(long long i=0;i<s.nn;i++){
            if(gmass) gmass[i]=G.mass[i];
            for(int a=0;a<d;a++){
                if(gvel) gvel[i*d+a]=G.velocity[i][a];
                if(gvsave) gvsave[i*d+a]=G.velocity_save[i][a];
                if(gvfric) gvfric[i*d+a]=G.velocity_friction_save[i][a];}}}
    return 0;
}
// Clamped plasticity with hardening (MPM_PLASTICITY_CLAMP, the snow model):
// the singular values of F are clamped to [1 - theta_c, 1 + theta_s]; when
// any changes, mu and lambda become mu0, lambda0 times
// exp(min(max_hardening, hardening_factor (1 - det Fp_new))) (at least 1e-8),
// F = U clamp(S) V^T and Fp = V clamp(S)^-1 U^T F Fp.
template<class T,int d> struct Plastic
{
    T theta_c,theta_s,max_hardening,hardening_factor;
    T* Fp;const T* mu0;const T* lambda0;T* mu;T* lambda;
};
template<class T> T Determinant(const Mat<T,3>& A)
{const T* x=A.x;return x[0]*(x[4]*x[8]-x[7]*x[5])+x[3]*(x[7]*x[2]-x[1]*x[8])+x[6]*(x[1]*x[5]-x[4]*x[2]);}
template<class T> T Determinant(const Mat<T,2>& A){const T* x=A.x;return x[0]*x[3]-x[1]*x[2];}

template<class T,int d> void Plasticity_Update(const Plastic<T,d>& pl,Particles<T,d>& P,const std::vector<unsigned char>& simulated)
{
    const T lo=1-pl.theta_c,hi=1+pl.theta_s;
    for(long long p=0;p<P.n;p++){
        if(!simulated[p]) continue;
        Mat<T,d> Fe=P.f(p),U,V;
        Vec<T,d> sv;
        Svd(Fe,U,sv,V);
        Vec<T,d> strain;
        bool same=true;
        for(int a=0;a<d;a++){
            strain[a]=sv[a]<=lo?lo:(sv[a]>=hi?hi:sv[a]);
            if(!(strain[a]==sv[a])) same=false;}
        if(same) continue;
        Mat<T,d> Fp;for(int k=0;k<d*d;k++) Fp.x[k]=pl.Fp[p*d*d+k];
        T det=Product(sv)/Product(strain)*Determinant(Fp);
        T hc=std::exp(std::min(pl.max_hardening,pl.hardening_factor*(1-det)));
        if(hc<(T)1e-8) hc=(T)1e-8;
        pl.mu[p]=pl.mu0[p]*hc;
        pl.lambda[p]=pl.lambda0[p]*hc;
        Mat<T,d> US;   // MATRIX * DIAGONAL_MATRIX: column j scaled
        for(int j=0;j<d;j++) for(int i=0;i<d;i++) US(i,j)=U(i,j)*strain[j];
        Mat<T,d> Fnew=Times_Transpose(US,V);
        Mat<T,d> DU;   // DIAGONAL_MATRIX::Times_Transpose: r(i,j)=1/s_i U(j,i)
        for(int i=0;i<d;i++) for(int j=0;j<d;j++) DU(i,j)=((T)1/strain[i])*U(j,i);
        Mat<T,d> Fpn=((V*DU)*Fe)*Fp;
        for(int k=0;k<d*d;k++){P.F[p*d*d+k]=Fnew.x[k];pl.Fp[p*d*d+k]=Fpn.x[k];}}
}

// Drucker-Prager (MPM_DRUCKER_PRAGER, sand): per particle plastic_def, rho_F
// (friction), sigma_Y (cohesion) and vc (volume correction).
template<class T,int d> struct DruckerPrager
{
    T a0,a1,a3,a4;bool use_vc;
    T* Fp;T* plastic_def;T* rho_F;T* sigma_Y;T* vc;
    void Update_Hardening(long long id,T inc) const
    {
        plastic_def[id]+=inc;
        T phi_F=M_PI/180*(a0+(a1*plastic_def[id]-a4)*std::exp(-a3*plastic_def[id]));
        T sin_phi_F=std::sin(phi_F);
        rho_F[id]=std::sqrt(2.0/3.0)*2*sin_phi_F/(3-sin_phi_F);
    }
    // MPM_DRUCKER_PRAGER::Compute with store_hardening
    bool Compute(Vec<T,d>& strain,const Vec<T,d>& Fe,long long id,T mu,T lambda) const
    {
        T g=1/(d*lambda+2*mu);
        T b=1/(2*mu);
        T beta=-sigma_Y[id]*g/(rho_F[id]*d);
        Vec<T,d> log_Fe;for(int a=0;a<d;a++) log_Fe[a]=std::log(std::abs(Fe[a]));
        Vec<T,d> strain_trial;
        T vcd=vc[id]/d;
        for(int a=0;a<d;a++) strain_trial[a]=(log_Fe[a]+beta)+vcd;
        T k=strain_trial[0];for(int a=1;a<d;a++) k+=strain_trial[a];
        Vec<T,d> sh;T kd=k/d;
        for(int a=0;a<d;a++) sh[a]=strain_trial[a]-kd;
        T q=magnitude(sh);
        T r=rho_F[id]*k/g;
        T dg=q+b*r;
        if(dg<=0) return false;
        if(q==0 || k>0){
            for(int a=0;a<d;a++) strain[a]=sigma_Y[id]?std::exp(-beta)*(T)1:(T)1;
            if(use_vc){T ls=log_Fe[0];for(int a=1;a<d;a++) ls+=log_Fe[a];vc[id]+=ls+beta*d;}
            Update_Hardening(id,magnitude(log_Fe));
            return true;}
        T p=r*b/q;
        for(int a=0;a<d;a++) strain[a]=std::exp((kd-p*sh[a])-beta);
        vc[id]=0;
        Update_Hardening(id,dg);
        return true;
    }
};

// MPM_PLASTICITY_MODEL::Update_Particles with the Drucker-Prager Compute.
template<class T,int d> void Drucker_Prager_Update(const DruckerPrager<T,d>& dp,Particles<T,d>& P,
    const std::vector<unsigned char>& simulated)
{
    for(long long p=0;p<P.n;p++){
        if(!simulated[p]) continue;
        Mat<T,d> Fe=P.f(p),U,V;
        Vec<T,d> sv;
        Svd(Fe,U,sv,V);
        Vec<T,d> strain;
        if(!dp.Compute(strain,sv,p,P.mu[p],P.lambda[p])) continue;
        Mat<T,d> Fp;for(int k=0;k<d*d;k++) Fp.x[k]=dp.Fp[p*d*d+k];
        Mat<T,d> US;
        for(int j=0;j<d;j++) for(int i=0;i<d;i++) US(i,j)=U(i,j)*strain[j];
        Mat<T,d> Fnew=Times_Transpose(US,V);
        Mat<T,d> DU;
        for(int i=0;i<d;i++) for(int j=0;j<d;j++) DU(i,j)=((T)1/strain[i])*U(j,i);
        Mat<T,d> Fpn=((V*DU)*Fe)*Fp;
        for(int k=0;k<d*d;k++){P.F[p*d*d+k]=Fnew.x[k];dp.Fp[p*d*d+k]=Fpn.x[k];}}
}

// state[4*np]: plastic_def, rho_F, sigma_Y, vc (initialized here from params
// a0, a1, a3, a4, sigma_Y0, use_vc as MPM_DRUCKER_PRAGER::Initialize_Particles).
template<class T,int d> int Run_Drucker_Prager(const int* N,const double* lo,const double* hi,const double* g,int model,
    double thr,double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi,
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,T* mu,T* la,unsigned char* valid,int nsteps,
    const double* params,T* Fp,T* state)
{
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi);
    DruckerPrager<T,d> dp{(T)params[0],(T)params[1],(T)params[2],(T)params[3],params[5]!=0,Fp,state,state+np,
        state+2*np,state+3*np};
    for(long long p=0;p<np;p++){
        dp.plastic_def[p]=0;dp.vc[p]=0;
        dp.sigma_Y[p]=(T)params[4];
        dp.Update_Hardening(p,0);}
    Particles<T,d> P{np,X,V,F,B,mass,volume,mu,la,valid};
    Grid<T,d> G;
    T time=0;
    std::vector<unsigned char> simulated(np);
    for(int k=0;k<nsteps;k++){
        for(long long p=0;p<np;p++) simulated[p]=valid[p];
        Step(s,P,G,time);
        Drucker_Prager_Update(dp,P,simulated);
        time+=s.dt;}
    return 0;
}

template<class T,int d> int Run_Plastic(const int* N,const double* lo,const double* hi,const double* g,int model,double thr,
    double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi,
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,T* mu,T* la,unsigned char* valid,int nsteps,
    const double* plastic,T* Fp)
{
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi);
    std::vector<T> mu0(mu,mu+np),la0(la,la+np);
    Plastic<T,d> pl{(T)plastic[0],(T)plastic[1],(T)plastic[2],(T)plastic[3],Fp,mu0.data(),la0.data(),mu,la};
    Particles<T,d> P{np,X,V,F,B,mass,volume,mu,la,valid};
    Grid<T,d> G;
    T time=0;
    std::vector<unsigned char> simulated(np);
    for(int k=0;k<nsteps;k++){
        for(long long p=0;p<np;p++) simulated[p]=valid[p];
        Step(s,P,G,time);
        Plasticity_Update(pl,P,simulated);
        time+=s.dt;}
    return 0;
}

// MPM_DRIVER::Grid_V_Upper_Bound (MPM_DRIVER.cpp:1125-1138) for use_affine and
// use_affine_cfl (MPM_EXAMPLE.h:116) with B stored: max over simulated
// particles of |v| + |B|_F * 6 sqrt(d) / min dx.
template<class T,int d> T Grid_V_Upper_Bound(const Scene<T,d>& s,const Particles<T,d>& P)
{
    T odx=s.one_over_dX[0];
    for(int a=1;a<d;a++) odx=std::min(odx,s.one_over_dX[a]);
    T xi=(T)6*std::sqrt((T)d)*odx;
    T result=0;
    for(long long p=0;p<P.n;p++){
        if(!P.valid[p]) continue;
        Vec<T,d> v;for(int a=0;a<d;a++) v[a]=P.V[p*d+a];
        Mat<T,d> B;for(int k=0;k<d*d;k++) B.x[k]=P.B[p*d*d+k];
        T r=magnitude(v)+std::sqrt(Frobenius_Norm_Squared(B))*xi;
        result=std::max(result,r);}
    return result;
}

// MPM_DRIVER::Compute_Dt (MPM_DRIVER.cpp:1098-1104).
template<class T,int d> T Compute_Dt(const Scene<T,d>& s,const Particles<T,d>& P,T cfl,T max_dt)
{
    T dxmin=s.dX[0];
    for(int a=1;a<d;a++) dxmin=std::min(dxmin,s.dX[a]);
    T critical_speed=cfl*dxmin/max_dt;
    T v=Grid_V_Upper_Bound(s,P);
    return (v>critical_speed)?(cfl*dxmin/v):max_dt;
}

// The substep loop of MPM_DRIVER::Simulate_To_Frame (MPM_DRIVER.cpp:173-197)
// for one frame.  When Reduce_Dt lowers a substep's dt, time advances by the
// dt taken and the frame goes on (:190-194).  Writes the substep dts taken.
template<class T,int d> int Advance_Frame(Scene<T,d>& s,Particles<T,d>& P,Grid<T,d>& G,T& time,T frame_dt,T cfl,
    T min_dt,T max_dt,T* dts,int max_sub)
{
    T time_at_frame=time+frame_dt;
    bool done=false;
    int n=0;
    for(int substep=0;!done;substep++){
        if(n>=max_sub) return -1;
        T dt=Compute_Dt(s,P,cfl,max_dt);
        dt=dt<=min_dt?min_dt:(dt>=max_dt?max_dt:dt);
        T next_time=time+dt;
        if(next_time+(T).0001*dt>time_at_frame){next_time=time_at_frame;done=true;}
        else if(next_time+dt>time_at_frame) next_time=(time+time_at_frame)/2;
        dt=next_time-time;
        s.dt=dt;
        T taken=Step(s,P,G,time);
        if(taken<dt){next_time=time+taken;done=false;}
        time=next_time;
        dts[n++]=taken;}
    return n;
}
// lim = {use_strong_cfl, use_sound_speed_cfl, cfl_F, cfl_sound} (null: both off).
template<class T,int d> void Set_Limits(Scene<T,d>& s,const double* lim,double cfl,double min_dt)
{
    s.cfl=(T)cfl;s.min_dt=(T)min_dt;
    if(!lim) return;
    s.strong_cfl=lim[0]!=0;s.sound_cfl=lim[1]!=0;s.cfl_F=(T)lim[2];s.cfl_sound=(T)lim[3];
}

// Fixed-dt steps with reflection boundary conditions.
template<class T,int d> int Run_Reflect(const int* N,const double* lo,const double* hi,const double* g,int model,double thr,
    double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi,
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps,
    const double* refl)
{
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi);
    s.reflect=1;s.full_reflect=refl[0]!=0;s.reflect_friction=(T)refl[1];
    for(int i=0;i<2*d;i++) s.side_type[i]=(int)refl[2+i];
    Particles<T,d> P{np,X,V,F,B,mass,volume,mu,la,valid};
    Grid<T,d> G;
    T time=0;
    for(int k=0;k<nsteps;k++) time+=Step(s,P,G,time);
    return 0;
}

// Fixed-dt steps with Reduce_Dt: dts[k] is the dt step k took.
template<class T,int d> int Run_Limited(const int* N,const double* lo,const double* hi,const double* g,int model,double thr,
    double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi,
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps,
    const double* lim,double cfl,double min_dt,double* dts)
{
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi);
    Set_Limits(s,lim,cfl,min_dt);
    Particles<T,d> P{np,X,V,F,B,mass,volume,mu,la,valid};
    Grid<T,d> G;
    T time=0;
    for(int k=0;k<nsteps;k++){T taken=Step(s,P,G,time);time+=taken;dts[k]=(double)taken;}
    return 0;
}
}  // namespace oracle

using namespace oracle;

#define DEFINE_FRAME(SUF,T) \
ORACLE_API_F int oracle_grid_v_bound_##SUF(int dim,const int* N,const double* lo,const double* hi,long long np,const T* V, \
    const T* B,const unsigned char* valid,double* out) \
{ \
    double g[3]={0,0,0}; \
    Particles<T,2> P2{np,0,(T*)V,0,(T*)B,0,0,0,0,(unsigned char*)valid}; \
    Particles<T,3> P3{np,0,(T*)V,0,(T*)B,0,0,0,0,(unsigned char*)valid}; \
    if(dim==2){*out=(double)Grid_V_Upper_Bound(Make_Scene<T,2>(N,lo,hi,g,0,0.25,1,0,0,0,0,0),P2);return 0;} \
    if(dim==3){*out=(double)Grid_V_Upper_Bound(Make_Scene<T,3>(N,lo,hi,g,0,0.25,1,0,0,0,0,0),P3);return 0;} \
    return -1; \
} \
ORACLE_API_F int oracle_advance_frames_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g,int model, \
    double thr,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid, \
    int nframes,double frame_dt,double cfl,double min_dt,double max_dt,double* dts,int max_sub,const double* lim) \
{ \
    int total=0; \
    if(dim==2){ \
        Scene<T,2> s=Make_Scene<T,2>(N,lo,hi,g,model,thr,max_dt,nwalls,wtype,wfric,wlo,whi); \
        Set_Limits(s,lim,cfl,min_dt); \
        Particles<T,2> P{np,X,V,F,B,mass,volume,mu,la,valid};Grid<T,2> G;T time=0;std::vector<T> d(max_sub); \
        for(int f=0;f<nframes;f++){int n=Advance_Frame(s,P,G,time,(T)frame_dt,(T)cfl,(T)min_dt,(T)max_dt,d.data(),max_sub-total); \
            if(n<0) return -1; \
            for(int i=0;i<n;i++) dts[total+i]=(double)d[i]; \
            total+=n;}} \
    else if(dim==3){ \
        Scene<T,3> s=Make_Scene<T,3>(N,lo,hi,g,model,thr,max_dt,nwalls,wtype,wfric,wlo,whi); \
        Set_Limits(s,lim,cfl,min_dt); \
        Particles<T,3> P{np,X,V,F,B,mass,volume,mu,la,valid};Grid<T,3> G;T time=0;std::vector<T> d(max_sub); \
        for(int f=0;f<nframes;f++){int n=Advance_Frame(s,P,G,time,(T)frame_dt,(T)cfl,(T)min_dt,(T)max_dt,d.data(),max_sub-total); \
            if(n<0) return -1; \
            for(int i=0;i<n;i++) dts[total+i]=(double)d[i]; \
            total+=n;}} \
    else return -1; \
    return total; \
}

#define ORACLE_API extern "C" __attribute__((visibility("default")))
#define ORACLE_API_F extern "C" __attribute__((visibility("default")))

#define DEFINE_RUN(SUF,T) \
ORACLE_API int oracle_run_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g,int model,double thr,double dt, \
    int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps, \
    T* gmass,T* gvel,T* gvsave,T* gvfric) \
{ \
    if(dim==2) return Run<T,2>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la,valid,nsteps,gmass,gvel,gvsave,gvfric); \
    if(dim==3) return Run<T,3>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la,valid,nsteps,gmass,gvel,gvsave,gvfric); \
    return -1; \
}
DEFINE_RUN(f64,double)
DEFINE_RUN(f32,float)

// Totals of a run, as -print_stats reports them each step: the particles
// after Update_Simulated_Particles, the grid after Particle_To_Grid (the
// velocity_save copy) and after Apply_Forces; then the particles at the end.
// out holds (3 nsteps + 1) records of 11 values (layout of Particle_Totals).
template<class T,int d> int Run_Totals(const int* N,const double* lo,const double* hi,const double* g,int model,double thr,
    double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi,
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps,
    T* out)
{
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi);
    Particles<T,d> P{np,X,V,F,B,mass,volume,mu,la,valid};
    Grid<T,d> G;
    T time=0;
    for(int k=0;k<nsteps;k++){
        Particle_Totals(s,P,out+(3*k)*11);
        time+=Step(s,P,G,time);
        Grid_Totals(s,G,G.velocity_save,out+(3*k+1)*11);
        Grid_Totals(s,G,G.velocity,out+(3*k+2)*11);}
    Particle_Totals(s,P,out+(3*nsteps)*11);
    return 0;
}
#define DEFINE_TOTALS(SUF,T) \
ORACLE_API int oracle_run_totals_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g,int model, \
    double thr,double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps, \
    T* out) \
{ \
    if(dim==2) return Run_Totals<T,2>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la,valid, \
        nsteps,out); \
    if(dim==3) return Run_Totals<T,3>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la,valid, \
        nsteps,out); \
    return -1; \
}
DEFINE_TOTALS(f64,double)
DEFINE_TOTALS(f32,float)
DEFINE_FRAME(f64,double)
DEFINE_FRAME(f32,float)

// plastic[4] = theta_c, theta_s, max_hardening, hardening_factor; mu0 and
// lambda0 are the mu and lambda passed in; Fp, mu, lambda are updated in place.
#define DEFINE_PLASTIC(SUF,T) \
ORACLE_API int oracle_run_plastic_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g,int model, \
    double thr,double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,T* mu,T* la,unsigned char* valid,int nsteps, \
    const double* plastic,T* Fp) \
{ \
    if(dim==2) return Run_Plastic<T,2>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la,valid, \
        nsteps,plastic,Fp); \
    if(dim==3) return Run_Plastic<T,3>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la,valid, \
        nsteps,plastic,Fp); \
    return -1; \
}
DEFINE_PLASTIC(f64,double)
DEFINE_PLASTIC(f32,float)

// params[6] = a0, a1, a3, a4, sigma_Y0, use_vc; state[4*np] out.
#define DEFINE_DP(SUF,T) \
ORACLE_API int oracle_run_drucker_prager_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g, \
    int model,double thr,double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,T* mu,T* la,unsigned char* valid,int nsteps, \
    const double* params,T* Fp,T* state) \
{ \
    if(dim==2) return Run_Drucker_Prager<T,2>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume, \
        mu,la,valid,nsteps,params,Fp,state); \
    if(dim==3) return Run_Drucker_Prager<T,3>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume, \
        mu,la,valid,nsteps,params,Fp,state); \
    return -1; \
}
DEFINE_DP(f64,double)
DEFINE_DP(f32,float)

// lim[4] = use_strong_cfl, use_sound_speed_cfl, cfl_F, cfl_sound.
#define DEFINE_LIMITED(SUF,T) \
ORACLE_API int oracle_run_limited_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g, \
    int model,double thr,double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps, \
    const double* lim,double cfl,double min_dt,double* dts) \
{ \
    if(dim==2) return Run_Limited<T,2>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la, \
        valid,nsteps,lim,cfl,min_dt,dts); \
    if(dim==3) return Run_Limited<T,3>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la, \
        valid,nsteps,lim,cfl,min_dt,dts); \
    return -1; \
}
DEFINE_LIMITED(f64,double)
DEFINE_LIMITED(f32,float)

// refl[2+2d] = use_full_reflection, reflection_bc_friction, side_bc_type per side.
#define DEFINE_REFLECT(SUF,T) \
ORACLE_API int oracle_run_reflect_##SUF(int dim,const int* N,const double* lo,const double* hi,const double* g, \
    int model,double thr,double dt,int nwalls,const int* wtype,const double* wfric,const double* wlo,const double* whi, \
    long long np,T* X,T* V,T* F,T* B,const T* mass,const T* volume,const T* mu,const T* la,unsigned char* valid,int nsteps, \
    const double* refl) \
{ \
    if(dim==2) return Run_Reflect<T,2>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la, \
        valid,nsteps,refl); \
    if(dim==3) return Run_Reflect<T,3>(N,lo,hi,g,model,thr,dt,nwalls,wtype,wfric,wlo,whi,np,X,V,F,B,mass,volume,mu,la, \
        valid,nsteps,refl); \
    return -1; \
}
DEFINE_REFLECT(f64,double)
DEFINE_REFLECT(f32,float)

// Sound_Speed of each particle from the singular values of its F (n x d*d).
#define DEFINE_SOUND(SUF,T) \
ORACLE_API int oracle_sound_speeds_##SUF(int dim,int model,double thr,long long n,const T* F,const T* mass, \
    const T* volume,const T* mu,const T* la,T* out) \
{ \
    for(long long p=0;p<n;p++){ \
        if(dim==2){Mat<T,2> A,U,V;Vec<T,2> s;for(int k=0;k<4;k++) A.x[k]=F[p*4+k];Svd(A,U,s,V); \
            out[p]=Sound_Speed<T,2>(model,s,mu[p],la[p],(T)thr,mass[p]/volume[p]);} \
        else if(dim==3){Mat<T,3> A,U,V;Vec<T,3> s;for(int k=0;k<9;k++) A.x[k]=F[p*9+k];Svd(A,U,s,V); \
            out[p]=Sound_Speed<T,3>(model,s,mu[p],la[p],(T)thr,mass[p]/volume[p]);} \
        else return -1;} \
    return 0; \
}
DEFINE_SOUND(f64,double)
DEFINE_SOUND(f32,float)

template<class T,int d> static void Svd_Many(long long n,const T* A,T* U,T* S,T* V,int* iters)
{
    for(long long i=0;i<n;i++){
        Mat<T,d> M,u,v;Vec<T,d> s;
        for(int k=0;k<d*d;k++) M.x[k]=A[i*d*d+k];
        int it=Svd(M,u,s,v);
        if(iters) iters[i]=it;
        for(int k=0;k<d*d;k++){U[i*d*d+k]=u.x[k];V[i*d*d+k]=v.x[k];}
        for(int k=0;k<d;k++) S[i*d+k]=s[k];}
}
#define DEFINE_SVD(SUF,T) \
ORACLE_API int oracle_svd_##SUF(int dim,long long n,const T* A,T* U,T* S,T* V,int* iters) \
{ \
    if(dim==2){Svd_Many<T,2>(n,A,U,S,V,iters);return 0;} \
    if(dim==3){Svd_Many<T,3>(n,A,U,S,V,iters);return 0;} \
    return -1; \
}
DEFINE_SVD(f64,double)
DEFINE_SVD(f32,float)

template<class T,int d> static void Pstrain_Many(int model,T thr,long long n,const T* S,const T* mu,const T* la,T* P)
{
    for(long long i=0;i<n;i++){
        Vec<T,d> s;for(int k=0;k<d;k++) s[k]=S[i*d+k];
        Vec<T,d> r=model==0?P_Corotated(s,mu[i],la[i]):(model==2?P_Hencky(s,mu[i],la[i]):P_Neo_Hookean(s,mu[i],la[i],thr));
        for(int k=0;k<d;k++) P[i*d+k]=r[k];}
}
#define DEFINE_PSTRAIN(SUF,T) \
ORACLE_API int oracle_pstrain_##SUF(int dim,int model,double thr,long long n,const T* S,const T* mu,const T* la,T* P) \
{ \
    if(dim==2){Pstrain_Many<T,2>(model,(T)thr,n,S,mu,la,P);return 0;} \
    if(dim==3){Pstrain_Many<T,3>(model,(T)thr,n,S,mu,la,P);return 0;} \
    return -1; \
}
DEFINE_PSTRAIN(f64,double)
DEFINE_PSTRAIN(f32,float)

// Kirchhoff-type stress V0*P(F)*F^T per particle (what the explicit step scatters).
template<class T,int d> static void Stress_Many(int model,T thr,long long n,const T* F,const T* vol,const T* mu,const T* la,T* A)
{
    int N[3]={1,1,1};double lo[3]={0,0,0},hi[3]={1,1,1},g[3]={0,0,0};
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,model,thr,1,0,0,0,0,0);
    for(long long i=0;i<n;i++){
        Mat<T,d> M;for(int k=0;k<d*d;k++) M.x[k]=F[i*d*d+k];
        Mat<T,d> R=vol[i]*Stress_PFT(s,M,M,mu[i],la[i]);
        for(int k=0;k<d*d;k++) A[i*d*d+k]=R.x[k];}
}
#define DEFINE_STRESS(SUF,T) \
ORACLE_API int oracle_stress_##SUF(int dim,int model,double thr,long long n,const T* F,const T* vol,const T* mu,const T* la,T* A) \
{ \
    if(dim==2){Stress_Many<T,2>(model,(T)thr,n,F,vol,mu,la,A);return 0;} \
    if(dim==3){Stress_Many<T,3>(model,(T)thr,n,F,vol,mu,la,A);return 0;} \
    return -1; \
}
DEFINE_STRESS(f64,double)
DEFINE_STRESS(f32,float)

template<class T,int d> static void Weights_Many(const int* N,const double* lo,const double* hi,long long n,const T* X,int* base,T* w,T* dw)
{
    double g[3]={0,0,0};
    Scene<T,d> s=Make_Scene<T,d>(N,lo,hi,g,0,0,1,0,0,0,0,0);
    for(long long i=0;i<n;i++){
        Vec<T,d> x;for(int a=0;a<d;a++) x[a]=X[i*d+a];
        Weights<T,d> pd;Compute_Weights(s,x,pd);
        for(int a=0;a<d;a++){
            base[i*d+a]=pd.base[a];
            for(int k=0;k<3;k++){w[(i*3+k)*d+a]=pd.w[k][a];dw[(i*3+k)*d+a]=pd.dw[k][a];}}}
}
#define DEFINE_WEIGHTS(SUF,T) \
ORACLE_API int oracle_weights_##SUF(int dim,const int* N,const double* lo,const double* hi,long long n,const T* X,int* base,T* w,T* dw) \
{ \
    if(dim==2){Weights_Many<T,2>(N,lo,hi,n,X,base,w,dw);return 0;} \
    if(dim==3){Weights_Many<T,3>(N,lo,hi,n,X,base,w,dw);return 0;} \
    return -1; \
}
DEFINE_WEIGHTS(f64,double)
DEFINE_WEIGHTS(f32,float)
