// The field API of src/field/field.h, source-compatible with the apps that
// include it (apps/euler/euler.cpp, apps/potential/potential.cpp), with
// every field resident in GPU memory.
//
//  * MeshField<T, E>: node values (CELL: every node of every cell, ghost
//    cells included, as the reference stores them; FACET: every face node)
//    in device memory, with a host copy made on demand for operator[] and
//    file I/O.  Assignments evaluate on the device.
//  * expressions (a + b, s * U, pow(x, y), mag(U), dot, mul, ...) build a
//    device functor tree; assigning one to a field launches one elementwise
//    kernel over the field (the kernel is instantiated where the app is
//    compiled with hipcc), with the reference's operation order.
//  * MeshMatrix, divf, gradf, cds, flx, lap, src, addTemporal, Solve,
//    applyExplicitBCs, fillBCs, reduce_*: the reference's operators, on the
//    device (operator kernels in field_ops.hip; the FV solve through the
//    library's PCG).
// The DG node and face-node numbering, tensor component order and file
// formats are the reference's.  Meshes with 2:1 mortar faces are refused.
#ifndef NSEM_FIELD_FIELD_H
#define NSEM_FIELD_FIELD_H

#include <cstdio>
#include <list>
#include <string>
#include <type_traits>
#include <vector>

#include "mp.h"
#include "tensor.h"
#include "util.h"

enum ENTITY { CELL, FACET, VERTEX, CELLMAT };
enum ACCESS { NO = 0, READ = 1, WRITE = 2, READWRITE = 3, STOREPREV = 4 };

// ---------------------------------------------------------------- device --
namespace nsem_field {
// The stream every field operation runs on, in order.
hipStream_t stream();
// Device memory from a size-keyed pool (the reference recycles its field
// buffers the same way); released only at exit.
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);
void upload(void* dst, const void* src, size_t bytes);
void download(void* dst, const void* src, size_t bytes);
void copy_d2d(void* dst, const void* src, size_t bytes);
void fail(const std::string& what);  // prints and exits, as the reference does
void check(hipError_t e, const char* what);
// element counts of the entities (set by Mesh::LoadMesh)
size_t entity_size(ENTITY e);
// reductions over the interior nodes of a cell field (ncomp doubles per
// node): kind 0 sum (per component), 1 max |v|, 2 min |v|
// sum (kind 0), max (1) or min (2) of the magnitudes over the first n
// values (a cell field's interior nodes, any other field's values)
void reduce(const Scalar* d, int ncomp, int kind, Scalar* out, size_t n);

template <class Op, class T, class A>
__global__ void assign_kernel(T* out, A a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Op::apply(out[i], a[i]);
}
struct AssignOp { template <class T, class V> NSEM_HD static void apply(T& o, const V& v) { o = v; } };
struct AddOp { template <class T, class V> NSEM_HD static void apply(T& o, const V& v) { o += v; } };
struct SubOp { template <class T, class V> NSEM_HD static void apply(T& o, const V& v) { o -= v; } };
struct MulOp { template <class T, class V> NSEM_HD static void apply(T& o, const V& v) { o *= v; } };
struct DivOp { template <class T, class V> NSEM_HD static void apply(T& o, const V& v) { o /= v; } };

template <class Op, class T, class A>
void launch_assign(T* out, const A& a, size_t n) {
    if (!n) return;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL((assign_kernel<Op, T, A>), dim3(blocks), dim3(256), 0, stream(), out, a, n);
    check(hipGetLastError(), "field assignment");
}

// expression leaves and nodes (device functors; copied by value into kernels)
template <class T>
struct Leaf {
    const T* p;
    NSEM_HD T operator[](size_t i) const { return p[i]; }
};
template <class T>
struct ConstLeaf {
    T v;
    NSEM_HD T operator[](size_t) const { return v; }
};
template <class F, class A>
struct Un {
    A a;
    NSEM_HD auto operator[](size_t i) const { return F::apply(a[i]); }
};
template <class F, class A, class B>
struct Bin {
    A a;
    B b;
    NSEM_HD auto operator[](size_t i) const { return F::apply(a[i], b[i]); }
};
#define NSEM_FN2(name, expr)                                                          \
    struct name {                                                                     \
        template <class X, class Y> NSEM_HD static auto apply(const X& x, const Y& y) { return expr; } \
    };
NSEM_FN2(FAdd, x + y)
NSEM_FN2(FSub, x - y)
NSEM_FN2(FMul, x * y)
NSEM_FN2(FDiv, x / y)
NSEM_FN2(FDot, dot(x, y))
NSEM_FN2(FOuter, mul(x, y))
NSEM_FN2(FPow, pow(x, y))
NSEM_FN2(FMax, max(x, y))
NSEM_FN2(FMin, min(x, y))
NSEM_FN2(FSdiv, sdiv(x, y))
NSEM_FN2(FDev, dev(x, y))
NSEM_FN2(FHyd, hyd(x, y))
#undef NSEM_FN2
#define NSEM_FN1(name, expr)                                                          \
    struct name {                                                                     \
        template <class X> NSEM_HD static auto apply(const X& x) { return expr; }     \
    };
NSEM_FN1(FNeg, -x)
NSEM_FN1(FMag, mag(x))
NSEM_FN1(FMagSq, magSq(x))
NSEM_FN1(FUnit, unit(x))
NSEM_FN1(FSqrt, sqrt(x))
NSEM_FN1(FExp, exp(x))
NSEM_FN1(FFabs, fabs(x))
NSEM_FN1(FTrn, trn(x))
NSEM_FN1(FSym, sym(x))
NSEM_FN1(FTr, tr(x))
#undef NSEM_FN1
// element-wise functions of Scalar values not in tensor.h
NSEM_HD Scalar sqrt(const Scalar& x) { return ::sqrt(x); }
NSEM_HD Scalar exp(const Scalar& x) { return nsem::dd::exp_cr(x); }
NSEM_HD Scalar pow(const Scalar& x, const Scalar& y) { return nsem::dd::pow_dd(x, y); }
NSEM_HD Scalar fabs(const Scalar& x) { return ::fabs(x); }
}  // namespace nsem_field

// An expression of element type T (the reference's DVExpr).
template <class T, class A>
struct DVExpr {
    A a;
};

// ------------------------------------------------------------ the fields --
class BaseField {
   public:
    std::string fName;
    Int fIndex = 0;
    ACCESS access = NO;
    static std::list<BaseField*> allFields;
    virtual ~BaseField() {}
    virtual void read_file(Int step) = 0;
    virtual void write_file(Int step) = 0;
    virtual int ncomponents() const = 0;
    virtual bool is_cell() const = 0;
};

namespace nsem_field {
// boundary conditions of one named field, per mesh boundary patch (the
// reference's BCondition list for that field, field.h:150-240)
enum BcType : int { BC_NONE = -1, BC_NEUMANN = 0, BC_SYMMETRY = 1, BC_DIRICHLET = 2, BC_ROBIN = 3, BC_CALC_DIRICHLET = 4 };
struct FieldBCs {
    std::vector<int> kind;         // per patch
    std::vector<Scalar> value;     // ncomp per patch
    std::vector<Scalar> shape;     // per patch (ROBIN)
    std::vector<Scalar> tvalue;    // ncomp per patch (ROBIN)
    std::vector<std::vector<Scalar>> fixed;  // CALC_DIRICHLET: per patch, ncomp per face node
};
FieldBCs& bcs_of(Int fIndex, int ncomp);
bool has_bcs(Int fIndex);
void read_field_file(const std::string& name, Int step, int ncomp, Scalar* host_all, Int fIndex);
void write_field_file(const std::string& name, Int step, int ncomp, const Scalar* host_all, Int fIndex);
// ghost nodes of a cell field from its boundary conditions
// (applyExplicitBCs, field.h:2585-2727)
void explicit_bcs(Scalar* d, int ncomp, Int fIndex);
// ghost node <- owner node; with bind, the gradient rule of the field bind
// (fillBCs, field.h:2730-2776); ncomp of the filled field, ncomp_src of the
// field whose conditions are used
void fill_bcs(Scalar* d, int ncomp, Int bind, int ncomp_src);
}  // namespace nsem_field

template <class type, ENTITY entity>
class MeshField : public BaseField {
    type* d_ = nullptr;           // device values
    mutable type* h_ = nullptr;   // host copy
    Int SIZE = 0;
    mutable bool host_valid_ = false;
    mutable bool dev_valid_ = true;

    void release() {
        if (d_) nsem_field::dev_free(d_, (size_t)SIZE * sizeof(type));
        delete[] h_;
        d_ = nullptr;
        h_ = nullptr;
    }

   public:
    static constexpr Int TYPE_SIZE = sizeof(type) / sizeof(Scalar);
    // previous time levels of the temporal schemes (tstore[0] newest)
    std::vector<MeshField*> tstore;
    Int nstored = 0;

    MeshField(const char* str = "", ACCESS a = NO, bool recycle = true) {
        (void)recycle;
        construct(str, a);
    }
    explicit MeshField(const bool) {}
    MeshField(const MeshField& p) {
        allocate();
        nsem_field::copy_d2d(d_, p.dptr(), (size_t)SIZE * sizeof(type));
    }
    MeshField(const type& v) {
        allocate();
        *this = v;
    }
    template <class A>
    MeshField(const DVExpr<type, A>& e) {
        allocate();
        nsem_field::launch_assign<nsem_field::AssignOp>(dptr_w(), e.a, SIZE);
    }
    ~MeshField() {
        if (fIndex) allFields.remove(this);
        for (auto* t : tstore) delete t;
        release();
    }
    void allocate() {
        release();
        SIZE = (Int)nsem_field::entity_size(entity);
        // the reference allocates one element more per CELL field
        const size_t n = (size_t)SIZE + (entity == CELL ? 1 : 0);
        d_ = static_cast<type*>(nsem_field::dev_alloc(n * sizeof(type)));
        host_valid_ = false;
        dev_valid_ = true;
    }
    void construct(const char* str = "", ACCESS a = NO) {
        access = a;
        fName = str;
        if (nsem_field::entity_size(entity)) allocate();
        fIndex = Util::hash_function(str);
        if (fIndex) allFields.push_back(this);
    }
    Int size() const { return SIZE; }

    // device pointers: read (uploading a newer host copy first) and write
    // (the host copy becomes stale)
    const type* dptr() const {
        if (!dev_valid_) {
            nsem_field::upload(d_, h_, (size_t)SIZE * sizeof(type));
            dev_valid_ = true;
        }
        return d_;
    }
    type* dptr_w() {
        dptr();
        host_valid_ = false;
        return d_;
    }
    // host access (as the reference's operator[], which may write): the
    // device copy is refreshed before the next device use
    type& operator[](Int i) const {
        if (!host_valid_) {
            if (!h_) h_ = new type[(size_t)SIZE + 1];
            nsem_field::download(h_, d_, (size_t)SIZE * sizeof(type));
            host_valid_ = true;
        }
        dev_valid_ = false;
        return h_[i];
    }
    const type* host() const {
        (*this)[0];
        return h_;
    }
    type* host_w() {
        (*this)[0];
        return h_;
    }

    // assignment
    MeshField& operator=(const MeshField& q) {
        if (this == &q) return *this;
        if (!d_) allocate();
        nsem_field::copy_d2d(dptr_w(), q.dptr(), (size_t)SIZE * sizeof(type));
        return *this;
    }
#define NSEM_MF_SOP(op, F)                                                                       \
    MeshField& operator op(const Scalar& v) {                                                    \
        if (!d_) allocate();                                                                     \
        nsem_field::launch_assign<nsem_field::F>(dptr_w(), nsem_field::ConstLeaf<Scalar>{v}, SIZE); \
        return *this;                                                                            \
    }
    NSEM_MF_SOP(+=, AddOp)
    NSEM_MF_SOP(-=, SubOp)
    NSEM_MF_SOP(*=, MulOp)
    NSEM_MF_SOP(/=, DivOp)
#undef NSEM_MF_SOP
    MeshField& operator=(const Scalar& v) {
        if (!d_) allocate();
        nsem_field::launch_assign<nsem_field::AssignOp>(dptr_w(), nsem_field::ConstLeaf<type>{type(v)}, SIZE);
        return *this;
    }
    template <int N, class X = type, class = std::enable_if_t<!std::is_same<X, Scalar>::value>>
    MeshField& operator=(const TTensor<N>& v) {
        if (!d_) allocate();
        nsem_field::launch_assign<nsem_field::AssignOp>(dptr_w(), nsem_field::ConstLeaf<type>{v}, SIZE);
        return *this;
    }
#define NSEM_MF_OP(op, F)                                                                        \
    MeshField& operator op(const MeshField& q) {                                                 \
        if (!d_) allocate();                                                                     \
        nsem_field::launch_assign<nsem_field::F>(dptr_w(), nsem_field::Leaf<type>{q.dptr()}, SIZE); \
        return *this;                                                                            \
    }                                                                                            \
    template <class A>                                                                           \
    MeshField& operator op(const DVExpr<type, A>& e) {                                           \
        if (!d_) allocate();                                                                     \
        nsem_field::launch_assign<nsem_field::F>(dptr_w(), e.a, SIZE);                           \
        return *this;                                                                            \
    }
    NSEM_MF_OP(+=, AddOp)
    NSEM_MF_OP(-=, SubOp)
    NSEM_MF_OP(*=, MulOp)
    NSEM_MF_OP(/=, DivOp)
#undef NSEM_MF_OP
    template <class A>
    MeshField& operator=(const DVExpr<type, A>& e) {
        if (!d_) allocate();
        nsem_field::launch_assign<nsem_field::AssignOp>(dptr_w(), e.a, SIZE);
        return *this;
    }
    // a field times a scalar field (e.g. r *= rho, as Relax and ddt use it)
    template <class X = type, class = std::enable_if_t<!std::is_same<X, Scalar>::value>>
    MeshField& operator*=(const MeshField<Scalar, entity>& q) {
        nsem_field::launch_assign<nsem_field::MulOp>(dptr_w(), nsem_field::Leaf<Scalar>{q.dptr()}, SIZE);
        return *this;
    }

    // time levels (field.h:1103-1116)
    void initStore(Int n, const MeshField& v) {
        for (auto* t : tstore) delete t;
        tstore.clear();
        for (Int k = 0; k < n; k++) tstore.push_back(new MeshField(v));
        nstored = 1;
        access = ACCESS(access | STOREPREV);
    }
    void updateStore(const MeshField& v) {
        for (Int k = (Int)tstore.size() - 1; k > 0; k--) *tstore[k] = *tstore[k - 1];
        if (!tstore.empty()) *tstore[0] = v;
        if (nstored < tstore.size()) nstored++;
    }

    // files (field.h:1407-1677): <name><step>.txt|.bin
    void read_file(Int step) override;
    void write_file(Int step) override;
    void write(Int step) { write_file(step); }
    int ncomponents() const override { return (int)TYPE_SIZE; }
    bool is_cell() const override { return entity == CELL; }
};

template <class type, ENTITY entity>
void MeshField<type, entity>::read_file(Int step) {
    if (!(access & READ) || entity != CELL) return;
    if (!d_) allocate();
    nsem_field::read_field_file(fName, step, (int)TYPE_SIZE, reinterpret_cast<Scalar*>(host_w()), fIndex);
}
template <class type, ENTITY entity>
void MeshField<type, entity>::write_file(Int step) {
    if (!(access & WRITE) || entity != CELL) return;
    nsem_field::write_field_file(fName, step, (int)TYPE_SIZE, reinterpret_cast<const Scalar*>(host()), fIndex);
}

typedef MeshField<Scalar, CELL> ScalarCellField;
typedef MeshField<Vector, CELL> VectorCellField;
typedef MeshField<STensor, CELL> STensorCellField;
typedef MeshField<Tensor, CELL> TensorCellField;
typedef MeshField<Scalar, FACET> ScalarFacetField;
typedef MeshField<Vector, FACET> VectorFacetField;
typedef MeshField<STensor, FACET> STensorFacetField;
typedef MeshField<Tensor, FACET> TensorFacetField;

// ------------------------------------------------------- expressions -----
namespace nsem_field {
template <class X> struct is_operand : std::false_type {};
template <class T, ENTITY E> struct is_operand<MeshField<T, E>> : std::true_type {};
template <class T, class A> struct is_operand<DVExpr<T, A>> : std::true_type {};
template <class X> struct is_constant : std::is_arithmetic<X> {};
template <int N> struct is_constant<TTensor<N>> : std::true_type {};

template <class T, ENTITY E> Leaf<T> node(const MeshField<T, E>& f) { return Leaf<T>{f.dptr()}; }
template <class T, class A> const A& node(const DVExpr<T, A>& e) { return e.a; }
template <class X, class = std::enable_if_t<std::is_arithmetic<X>::value>>
ConstLeaf<Scalar> node(const X& v) { return ConstLeaf<Scalar>{Scalar(v)}; }
template <int N> ConstLeaf<TTensor<N>> node(const TTensor<N>& v) { return ConstLeaf<TTensor<N>>{v}; }

template <class X> using node_t = std::decay_t<decltype(node(std::declval<const X&>()))>;
template <class A> using value_t = std::decay_t<decltype(std::declval<const A&>()[0])>;

template <class F, class L, class R>
auto make_bin(const L& l, const R& r) {
    using B = Bin<F, node_t<L>, node_t<R>>;
    return DVExpr<value_t<B>, B>{B{node(l), node(r)}};
}
template <class F, class X>
auto make_un(const X& x) {
    using U = Un<F, node_t<X>>;
    return DVExpr<value_t<U>, U>{U{node(x)}};
}
template <class L, class R>
using binary_ok = std::enable_if_t<(is_operand<L>::value || is_operand<R>::value) &&
                                   (is_operand<L>::value || is_constant<L>::value) &&
                                   (is_operand<R>::value || is_constant<R>::value)>;
template <class X>
using unary_ok = std::enable_if_t<is_operand<X>::value>;
}  // namespace nsem_field

#define NSEM_EXPR_BIN(op, F)                                                          \
    template <class L, class R, class = nsem_field::binary_ok<L, R>>                  \
    auto operator op(const L& l, const R& r) {                                        \
        return nsem_field::make_bin<nsem_field::F>(l, r);                             \
    }
NSEM_EXPR_BIN(+, FAdd)
NSEM_EXPR_BIN(-, FSub)
NSEM_EXPR_BIN(*, FMul)
NSEM_EXPR_BIN(/, FDiv)
#undef NSEM_EXPR_BIN
#define NSEM_EXPR_FN2(name, F)                                                        \
    template <class L, class R, class = nsem_field::binary_ok<L, R>>                  \
    auto name(const L& l, const R& r) {                                               \
        return nsem_field::make_bin<nsem_field::F>(l, r);                             \
    }
NSEM_EXPR_FN2(dot, FDot)
NSEM_EXPR_FN2(mul, FOuter)
NSEM_EXPR_FN2(pow, FPow)
NSEM_EXPR_FN2(max, FMax)
NSEM_EXPR_FN2(min, FMin)
NSEM_EXPR_FN2(sdiv, FSdiv)
NSEM_EXPR_FN2(dev, FDev)
NSEM_EXPR_FN2(hyd, FHyd)
#undef NSEM_EXPR_FN2
#define NSEM_EXPR_FN1(name, F)                                                        \
    template <class X, class = nsem_field::unary_ok<X>>                               \
    auto name(const X& x) {                                                           \
        return nsem_field::make_un<nsem_field::F>(x);                                 \
    }
NSEM_EXPR_FN1(operator-, FNeg)
NSEM_EXPR_FN1(mag, FMag)
NSEM_EXPR_FN1(magSq, FMagSq)
NSEM_EXPR_FN1(unit, FUnit)
NSEM_EXPR_FN1(sqrt, FSqrt)
NSEM_EXPR_FN1(exp, FExp)
NSEM_EXPR_FN1(fabs, FFabs)
NSEM_EXPR_FN1(trn, FTrn)
NSEM_EXPR_FN1(sym, FSym)
NSEM_EXPR_FN1(tr, FTr)
#undef NSEM_EXPR_FN1

// an expression evaluated into a new cell field
template <class T, class A>
MeshField<T, CELL> eval_expr(const DVExpr<T, A>& e) {
    return MeshField<T, CELL>(e);
}
template <class T, ENTITY E>
const MeshField<T, E>& eval_expr(const MeshField<T, E>& f) {
    return f;
}

// ------------------------------------------------------------ reductions --
namespace Mesh {
extern Int gBCSfield;
}
// the reductions run over a cell field's interior nodes (gBCSfield) and
// over every value of other fields
template <class T, ENTITY E>
size_t reduce_count(const MeshField<T, E>& p) {
    return E == CELL ? (size_t)Mesh::gBCSfield : (size_t)p.size();
}
template <class T, ENTITY E>
T reduce_sum(const MeshField<T, E>& p, bool = false) {
    T s;
    nsem_field::reduce(reinterpret_cast<const Scalar*>(p.dptr()), (int)MeshField<T, E>::TYPE_SIZE, 0,
                       reinterpret_cast<Scalar*>(&s), reduce_count(p));
    return s;
}
template <class T, ENTITY E>
Scalar reduce_max(const MeshField<T, E>& p, bool = false) {
    Scalar s;
    nsem_field::reduce(reinterpret_cast<const Scalar*>(p.dptr()), (int)MeshField<T, E>::TYPE_SIZE, 1, &s, reduce_count(p));
    return s;
}
template <class T, ENTITY E>
Scalar reduce_min(const MeshField<T, E>& p, bool = false) {
    Scalar s;
    nsem_field::reduce(reinterpret_cast<const Scalar*>(p.dptr()), (int)MeshField<T, E>::TYPE_SIZE, 2, &s, reduce_count(p));
    return s;
}
template <class T, class A>
T reduce_sum(const DVExpr<T, A>& e) {
    return reduce_sum(MeshField<T, CELL>(e));
}

// ---------------------------------------------------- global namespaces --
namespace DG {
extern Int Nop[3];
extern Int NPX, NPY, NPZ, NP, NPF, NPMAT;
}  // namespace DG

namespace Controls {
enum FILE_FORMAT { TEXT = 0, BINARY = 1 };
enum Scheme { CDS, UDS, BLENDED, HYBRID, RUSANOV, LUD, CDSS, MUSCL, QUICK, VANLEER, VANALBADA, MINMOD,
              SUPERBEE, SWEBY, QUICKL, UMIST, DDS, FROMM };
enum NonOrthoScheme { NONE, MINIMUM, ORTHOGONAL, OVER_RELAXED };
enum TimeScheme { BDF1, BDF2, BDF3, BDF4, BDF5, BDF6, AM1, AM2, AM3, AM4, AM5, AB1, AB2, AB3, AB4, AB5,
                  RK1, RK2, RK3, RK4 };
enum Solvers { JACOBI, SOR, PCG };
enum Preconditioners { NOPR, DIAG, SSOR, DILU };
enum CommMethod { BLOCKED, ASYNCHRONOUS };
enum State { STEADY, TRANSIENT };
extern Scheme convection_scheme;
extern Int TVDbruner;
extern NonOrthoScheme nonortho_scheme;
extern TimeScheme time_scheme;
extern Solvers Solver;
extern Preconditioners Preconditioner;
extern CommMethod parallel_method;
extern State state;
extern Scalar SOR_omega, tolerance, blend_factor, dt;
extern Int max_iterations, write_interval, start_step, end_step, amr_step, current_step, n_deferred,
    save_average, print_time;
extern Vector gravity;
extern FILE_FORMAT write_format;
void enrollRefine(Util::ParamList& params);
void enrollDecompose(Util::ParamList& params);
}  // namespace Controls

namespace Mesh {
extern std::string gMeshName;
extern Int gBCS, gNCells, gNFacets, gBCSfield, gALLfield;
extern Int is_spherical;
extern Scalar sphere_radius, sphere_height;
extern VectorCellField cC;
extern ScalarCellField cV;
void enroll(Util::ParamList& params);
void LoadMesh(Int step = 0);
Vector calc_courant(const VectorCellField& U, Scalar dt);
void read_fields(Int step);
void write_fields(Int step);
template <class type>
void fixedBCs(const MeshField<type, CELL>& src, MeshField<type, CELL>& dest);
template <class type>
void scaleBCs(const MeshField<type, CELL>& src, MeshField<type, CELL>& dest, Scalar psi);
}  // namespace Mesh

namespace Prepare {
void refineMesh(Int step);
}

namespace nsem_field {
// boundary conditions copied from one field to another: CALC_DIRICHLET
// (fixedBCs, field.h:2778-2794) or scaled (scaleBCs, field.h:2796-2826)
void copy_bcs(Int src, Int dst, int ncomp, int mode, Scalar psi);
}  // namespace nsem_field

template <class type>
void Mesh::fixedBCs(const MeshField<type, CELL>& src, MeshField<type, CELL>& dest) {
    nsem_field::copy_bcs(src.fIndex, dest.fIndex, (int)MeshField<type, CELL>::TYPE_SIZE, 0, 0.0);
}
template <class type>
void Mesh::scaleBCs(const MeshField<type, CELL>& src, MeshField<type, CELL>& dest, Scalar psi) {
    nsem_field::copy_bcs(src.fIndex, dest.fIndex, (int)MeshField<type, CELL>::TYPE_SIZE, 1, psi);
}

// ------------------------------------------------------ boundary passes --
template <class T, ENTITY E>
void applyExplicitBCs(const MeshField<T, E>& cF, bool = false) {
    if (E != CELL) return;
    MeshField<T, E>& f = const_cast<MeshField<T, E>&>(cF);
    nsem_field::explicit_bcs(reinterpret_cast<Scalar*>(f.dptr_w()), (int)MeshField<T, E>::TYPE_SIZE, f.fIndex);
}
template <class T>
void fillBCs(MeshField<T, CELL>& cF, Int bind = 0, bool = false) {
    nsem_field::fill_bcs(reinterpret_cast<Scalar*>(cF.dptr_w()), (int)MeshField<T, CELL>::TYPE_SIZE, bind,
                         (int)MeshField<T, CELL>::TYPE_SIZE);
}

#include "field_ops.h"

#endif
