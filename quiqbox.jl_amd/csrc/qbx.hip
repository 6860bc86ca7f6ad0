// qbx.hip -- libqbx_hip.so: Gaussian integrals and direct Fock build for
// MI355X (gfx950), behind the C-ABI of include/qbx.h.
//
// Layout in HBM (one basis):
//   DevShell[n_shell]            centre, L, component -> basis function map
//   PrimPair[n_primpair]         primitive-pair records of every shell pair
//   per pair class (la >= lb):   oriented shell pairs sorted by Schwarz bound
//                                (descending): shell ids, primitive-pair
//                                range, Q, A-B
//   per ERI class and precision: int64 prefix over bra pairs of the number of
//                                surviving kets, and the first ket of each bra
//   D (2*n_set*N^2), J, K, G (n_set*N^2) FP64 work buffers
//
// One thread evaluates one contracted shell quartet: primitive loop with the
// generated vertical recurrence, contraction, generated horizontal recurrence,
// then the J/K digest with per-thread pre-reduction and FP64 atomics.
#include "qbx.h"
#include "qbx_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "qbx_common.h"
#include "qbx_eri_api.h"
#include "qbx_linalg.h"
#include "gen/eri_table.inc"

using namespace qbx;

namespace {
using namespace qbx;

using qbx::QbxError;

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t cnt)
    {
        if (cnt <= n && p) return;
        release();
        size_t c = cnt ? cnt : 1;
        hipError_t e = hipMalloc((void**)&p, c * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw QbxError(QBX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        n = c;
    }
    void upload(const std::vector<T>& v, hipStream_t s)
    {
        ensure(v.size());
        if (!v.empty()) QBX_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

constexpr double kPi = 3.14159265358979323846;
static_assert(QBX_MAXC == QBX_SHELL_MAXC && QBX_MAXL == QBX_MAX_L, "shell limits of qbx.h and the kernels differ");
static_assert(QBX_PAIR_LMAX == QBX_MAXL, "pair classes must cover every shell");
// auxiliary streams of a Fock build: a process gets 4 hardware queues, one of
// them the context's main stream, so a fourth auxiliary stream would share a
// queue with another and leave its work serialized behind it
constexpr int kStreams = 3;
// wave slots of a staged class's workspace (persistent grid): 8 per CU
constexpr int kStageSlots = 2048;
// primitive pairs with |c_a c_b exp(-ab/p R^2) / p| below this are dropped.
// Each primitive quartet the cut removes is below 1e-20 times the other
// pair's factor (O(1e2) at most), the Boys value (at most 1) and the
// polynomial factors of the recurrences.  Measured against no cut: the Fock
// matrices of the 592-function s/p/d cluster and of 24 H2O in cc-pVDZ move
// by at most 3e-12, their energies by 3e-11 Eh; the water build has 20%
// fewer primitive quartets.
constexpr double QBX_PRIM_PAIR_CUT = 1e-20;
// Ket buckets.  A wave's lanes take consecutive kets of one bra and the wave
// runs as long as its lane with the most primitive pairs; in Schwarz order,
// contracted core pairs (81 primitive pairs in cc-pVDZ oxygen) sit among
// single-primitive ones.  Each pair class is therefore ordered by bucket of
// primitive-pair count, Schwarz order within a bucket, and every bra gets one
// ket range per bucket.  Estimated on the (ss|ss) kets of 24 H2O in cc-pVDZ:
// lanes busy 36% -> 78% of the primitive loop, 3.5% more waves.
constexpr int kKetBuckets = 5;

// Fixed device work array of qbx_density_purify_noalloc: P, Q, T (n x n) and
// 2n + 4 auxiliary values for n up to QBX_PURIFY_NOALLOC_MAX_N, 25 MB per
// GPU for the life of the process.  Host code never uses the symbol itself as
// an address: qbx_init takes each device's address with hipGetSymbolAddress.
constexpr size_t kPurifyScratch = 3 * (size_t)QBX_PURIFY_NOALLOC_MAX_N * QBX_PURIFY_NOALLOC_MAX_N +
                                  2 * (size_t)QBX_PURIFY_NOALLOC_MAX_N + 4;
static __device__ __align__(256) double scratch[kPurifyScratch];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

inline int ket_bucket(int ppn) { return ppn <= 1 ? 0 : ppn <= 4 ? 1 : ppn <= 9 ? 2 : ppn <= 27 ? 3 : 4; }

}  // namespace

// ---------------------------------------------------------------------------
// context and basis objects
// ---------------------------------------------------------------------------
struct qbx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[kStreams] = {};   // concurrent ERI-class launches (default priority)
    hipEvent_t ev_fork = nullptr, ev_join[kStreams] = {};
    DevBuf<double> boys, bexp;
    DevBuf<double> boysl;          // per-order rows for the ERI kernels (LDS-staged)
    int timing = 0;                // 0 off, 1 serialized per-class timing, 2 concurrent per-class timing
    std::vector<float> last_ms;
    float last_total = 0.f;
    // concurrent timing: a ring of event sets, one per Fock build, read back
    // (and summed) when a set is reused or at qbx_last_timing
    struct TimingSet {
        hipEvent_t beg_l[QBX_N_CLASSES], end_l[QBX_N_CLASSES], beg_s[QBX_N_CLASSES], end_s[QBX_N_CLASSES];
        hipEvent_t b0, b1;
        bool pending;
    };
    std::vector<TimingSet> tsets;
    int tnext = 0;
    std::vector<double> tsum;      // summed per-class milliseconds of the drained sets
    double tsum_total = 0.0;
    int tcount = 0;
    double mixed_cut = 1e-5;
    double density_thr = 0.0;      // > 0: density-screened J/K
    bool xcd_local = false;        // per-XCD J/K copies (set by the probe at init)
    int n_xcd = 1;
    int lds_max = 65536;           // LDS bytes per block on this device
    bool twins = true;             // twin-bra groups in the J/K build (QBX_TWINS=0: off)
    bool ket_buckets = true;       // kets bucketed by primitive-pair count (QBX_KET_BUCKETS=0: off)
    bool jk_rows = true;           // LDS K-row and twin kernels in the J/K build (QBX_JK_ROWS=0: off, the
                                   // one-wave-per-chunk kernels run everywhere)
    bool staged_split = true;      // (dd|dd), (dd|dp), (dp|dp): 64-ket chunks shared by a block's waves (QBX_STAGED_SPLIT=0: off)
    DevBuf<double> la_X, la_F, la_P, la_Q, la_T, la_aux;   // density purification
    double* scratch = nullptr;     // this device's address of the fixed work array
    rocblas_handle blas = nullptr; // rocSOLVER's handle (qbx_fock_solve), created on first use
    std::mutex mu;
};

namespace {
struct PairClass {
    // host copies (once Schwarz bounds exist: bucket-major, Q descending within
    // a bucket; rank = position in the plain Q-descending order, bstart = first
    // position of each ket bucket)
    std::vector<int> shA, shB, ppo, ppn;
    std::vector<double> Q, AB;
    std::vector<int> rank, bstart;
    int nbk = 1;               // ket buckets of this class (1: plain Schwarz order)
    std::unique_ptr<PairClass> plain;   // bucketed classes: the same pairs in plain Schwarz order
    DevBuf<int> d_shA, d_shB, d_ppo, d_ppn;
    DevBuf<double> d_Q, d_AB;
    int n() const { return (int)shA.size(); }
};

struct TaskSet {
    DevBuf<int64_t> prefix;
    DevBuf<int> koff;
    DevBuf<int> cpre;          // waves (chunks of 64 kets) per bra, prefix sums
    DevBuf<int> spre;          // K-row blocks (chunks chunks) per bra, prefix sums
    int nsuper = 0;
    int chunks = ROW_CHUNKS;   // chunks per K-row task
    int64_t total = 0;
    int nwave = 0;
    int nbra = 0;
};

// Twin-bra groups of one ERI class and precision (see TWIN_MAX): task prefix
// and, per group, TWIN_MAX records (bra, first ket, end of kets).
struct TwinSet {
    DevBuf<int> gpre;
    DevBuf<int> gmem;
    int ngroup = 0, ntask = 0, twmax = 0;
};

struct HostShell {
    double c[3];
    int L, nprim, poff;
    int bf[QBX_MAXC];
    double sc[QBX_MAXC];
};
}  // namespace

struct qbx_basis {
    qbx_ctx* ctx = nullptr;
    int nbf = 0, maxL = 0;
    std::vector<HostShell> shells;
    std::vector<double> pexp, pcoef;
    std::vector<PrimPair> pp;
    std::vector<int> ppk;           // radial indices of each primitive pair: k_a | k_b << 16
    PairClass pc[QBX_N_PAIR_CLASSES];
    DevBuf<PrimPair> d_pp;
    DevBuf<int> d_ppk, d_poff;
    DevBuf<DevShell> d_sh;
    // all shell pairs for the one-electron kernels (index into pc lists)
    std::vector<int2> p1;       // (class, index)
    DevBuf<int> d_p1A, d_p1B, d_p1o, d_p1n;
    bool schwarz = false;
    // task cache
    bool tasks_valid = false;
    double task_thr = -1.0, task_cut = -1.0;
    int task_mode = -1;
    TaskSet t64[QBX_N_CLASSES], t32[QBX_N_CLASSES];
    double class_cost[QBX_N_CLASSES] = {};   // estimated work of each class (launch order)
    // twin bras: groups of positions in each (Schwarz-sorted) pair class; for
    // the J/K build of a class with groups, the other bras' task lists and
    // the groups' own
    std::vector<std::vector<int>> twins[QBX_N_PAIR_CLASSES];
    bool use_twin[QBX_N_CLASSES] = {};
    TaskSet s64[QBX_N_CLASSES], s32[QBX_N_CLASSES];
    TwinSet w64[QBX_N_CLASSES];
    DevBuf<double> ws[QBX_N_CLASSES];   // staged-class workspaces
    // work buffers
    DevBuf<double> d_D, d_J, d_K, d_G;
    DevBuf<double> d_pdmax;        // shell-pair max |D| for density screening
    // end of the last Fock build (recorded on its stream after the
    // symmetrization) and that stream: a build on another stream, and every
    // rewrite of the task lists, waits for it before it touches the shared
    // buffers (D, J, K, task lists, workspaces)
    hipEvent_t ev_done = nullptr;
    hipStream_t last_stream = nullptr;
    bool built = false;
    // shell quartets with a shell above the generated classes (f, g): the
    // generic kernel's bra and ket lists at the threshold `thr`
    bool has_high = false;
    int nm = 0;                     // its recurrence size, 4 x the highest L
    DevBuf<double> d_pexp;
    struct HighLists {
        bool valid = false;
        double thr = -1.0;
        int nbra = 0;
        int64_t total = 0;
        DevBuf<int> hA, hB, ho, hn, lA, lB, lo, ln, nlow;
        DevBuf<double> hQ, lQ;
        DevBuf<int64_t> prefix;
    } hl;
    ~qbx_basis()
    {
        if (ev_done) (void)hipEventDestroy(ev_done);
    }
};

namespace {

// ---------------------------------------------------------------------------
// Boys table (host): F_m(T_i) by the convergent series
//   F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
// at the top order, then the stable downward recurrence.
// ---------------------------------------------------------------------------
// F_0..F_mtop at T by the convergent series at the top order and the stable
// downward recurrence, in long double; returns e^{-T}.
long double boys_seq_ld(long double T, int mtop, long double* F)
{
    const long double e = expl(-T);
    long double term = 1.0L / (2 * mtop + 1), sum = term;
    for (int k = 1; k < 20000; ++k) {
        term *= 2.0L * T / (2 * mtop + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
    }
    F[mtop] = e * sum;
    for (int m = mtop - 1; m >= 0; --m) F[m] = (2.0L * T * F[m + 1] + e) / (2 * m + 1);
    return e;
}

void build_boys_table(std::vector<double>& tab, std::vector<double>& et, std::vector<double>& rows)
{
    static_assert(QBX_BOYSL_LMAX + QBX_BOYSL_TAYLOR <= 24, "Boys row orders too high");
    static_assert((QBX_BOYSL_ROWS - 2) * QBX_BOYSL_DT >= QBX_BOYS_TMAX, "Boys rows do not cover the table range");
    tab.assign((size_t)QBX_BOYS_ROWS * QBX_BOYS_NCOL, 0.0);
    et.assign(QBX_BOYS_ROWS, 0.0);
    rows.assign((size_t)(QBX_BOYSL_LMAX + 1) * QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE, 0.0);
    long double F[64];
    static_assert(QBX_BOYS_NCOL <= 64, "one-electron Boys table orders too high");
    // one-electron kernels, the primitive kernels and the generic ERI kernel: F_0..F_32 on the 0.1 grid
    for (int i = 0; i < QBX_BOYS_ROWS; ++i) {
        const long double e = boys_seq_ld((long double)i * QBX_BOYS_DT, QBX_BOYS_NCOL - 1, F);
        for (int m = 0; m < QBX_BOYS_NCOL; ++m) tab[(size_t)i * QBX_BOYS_NCOL + m] = (double)F[m];
        et[i] = (double)e;
    }
    // ERI kernels: per-order Taylor rows on the 0.25 grid
    for (int i = 0; i < QBX_BOYSL_ROWS; ++i) {
        const long double e = boys_seq_ld((long double)i * QBX_BOYSL_DT, 23, F);
        for (int L = 0; L <= QBX_BOYSL_LMAX; ++L) {
            double* r = &rows[((size_t)L * QBX_BOYSL_ROWS + i) * QBX_BOYSL_STRIDE];
            long double fact = 1.0L;
            for (int k = 0; k < QBX_BOYSL_TAYLOR; ++k) {
                if (k > 0) fact *= k;
                r[k] = (double)(F[L + k] / fact);
            }
            r[QBX_BOYSL_TAYLOR] = (double)e;
        }
    }
}

// ---------------------------------------------------------------------------
// kernels: one-electron integrals
// ---------------------------------------------------------------------------
enum OneE { ONE_OVERLAP = 0, ONE_KINETIC = 1, ONE_NUCLEAR = 2, ONE_MULTIPOLE = 3, ONE_NUCLEAR_OWN = 4 };

struct OneArgs {
    const DevShell* sh;
    const PrimPair* pp;
    const int* pA;
    const int* pB;
    const int* po;
    const int* pn;
    int npair;
    int N;
    const double* boys;
    const double* bexp;
    int nnuc;
    const double* Z;
    const double* R;
    const int* own;        // ONE_NUCLEAR_OWN: nucleus at each shell's centre (-1: none)
    double M[3];
    int deg[3];
    int mc;                // Cartesian components per shell in the thread layout
    double* out;
};

// Nuclear-attraction primitive value for one component pair and one point C:
// the Obara-Saika one-body scheme axis by axis with the Boys auxiliary order
// carried from axis to axis (the scheme of GaussianOrbitals.jl:478-522).
template <int NMAX>
__device__ double coulomb1_comp_t(const int* la, const int* lb, const double* PA, const double* PC,
                                  const double* AB, double inv2p, const double* F, int Ltot)
{
    double vals[NMAX + 1];
    for (int m = 0; m <= Ltot; ++m) vals[m] = F[m];
    int nUp = Ltot;
    for (int t = 0; t < 3; ++t) {
        const int a = la[t], b = lb[t], S = a + b;
        if (S == 0) continue;
        double T[NMAX + 1][NMAX + 1];
        for (int j = 0; j <= nUp; ++j) T[0][j] = vals[j];
        for (int i = 0; i < S; ++i)
            for (int j = 0; j <= nUp - i - 1; ++j) {
                double v = PA[t] * T[i][j] - PC[t] * T[i][j + 1];
                if (i > 0) v += (i * inv2p) * (T[i - 1][j] - T[i - 1][j + 1]);
                T[i + 1][j] = v;
            }
        for (int m = 0; m <= nUp - S; ++m) {
            double h[NMAX + 1];
            for (int i = a; i <= S; ++i) h[i] = T[i][m];
            for (int s = 0; s < b; ++s)
                for (int k = a; k < a + b - s; ++k) h[k] = h[k + 1] + AB[t] * h[k];
            vals[m] = h[a];
        }
        nUp -= S;
    }
    return vals[0];
}


// ---------------------------------------------------------------------------
// kernel: one-electron integrals of primitive pairs (the reference's
// per-primitive engines), any angular momentum up to kPrimLMax per primitive
// ---------------------------------------------------------------------------
constexpr int kPrimLMax = 10;
constexpr int kPrimOv = kPrimLMax + 5;   // per-axis tables: j up to l_b + 4 (multipole degree) or l_b + 2 (kinetic)

struct PrimOneArgs {
    int n;
    int kind;                  // ONE_OVERLAP, ONE_KINETIC, ONE_NUCLEAR (1/|r - C|, no charge), ONE_MULTIPOLE
    const double* A;           // [n][3] centres and exponents of the first primitives
    const double* a;
    const int* la;             // [n][3] Cartesian exponents
    const double* B;
    const double* b;
    const int* lb;
    double C[3];
    int deg[3];
    const double* boys;
    const double* bexp;
    double* out;
};

// F_0..F_n for any T >= 0 and n <= QBX_BOYS_NMAX: the table below its range;
// above it F_0 = sqrt(pi/4T) and the upward recurrence with its e^{-T} term,
// F_{m+1} = ((2m+1) F_m - e^{-T}) / 2T (stable for T > m).
__device__ void prim_boys(int n, double T, double* F, const double* boys, const double* bexp)
{
    if (T < QBX_BOYS_TMAX) {
        qbx_boys_rt(n, T, F, boys, bexp);
        return;
    }
    const double e = exp(-T), r = 0.5 / T;
    F[0] = 0.88622692545275801365 * rsqrt(T);
    for (int m = 0; m < n; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * r;
}

__global__ __launch_bounds__(256) void prim_one_e_kernel(PrimOneArgs g)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= g.n) return;
    const double a = g.a[i], b = g.b[i], p = a + b, inv2p = 0.5 / p;
    int la[3], lb[3];
    double A[3], B[3], AB[3], PA[3], PB[3];
    double r2 = 0.0;
    for (int t = 0; t < 3; ++t) {
        la[t] = g.la[3 * i + t];
        lb[t] = g.lb[3 * i + t];
        A[t] = g.A[3 * i + t];
        B[t] = g.B[3 * i + t];
        AB[t] = A[t] - B[t];
        r2 += AB[t] * AB[t];
        const double P = (a * A[t] + b * B[t]) / p;
        PA[t] = P - A[t];
        PB[t] = P - B[t];
    }
    const double kab = exp(-a * b / p * r2);
    double val;
    if (g.kind == ONE_NUCLEAR) {
        const int Lt = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2];
        double PC[3];
        for (int t = 0; t < 3; ++t) PC[t] = PA[t] + A[t] - g.C[t];
        const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        double F[2 * kPrimLMax + 1];
        prim_boys(Lt, T, F, g.boys, g.bexp);
        val = 2.0 * kPi / p * kab * coulomb1_comp_t<2 * kPrimLMax>(la, lb, PA, PC, AB, inv2p, F, Lt);
    } else {
        const int extra = g.kind == ONE_KINETIC ? 2 : 0;
        double f[3], s0[3];
        for (int t = 0; t < 3; ++t) {
            const int nj = lb[t] + (g.kind == ONE_MULTIPOLE ? g.deg[t] : extra);
            double S[kPrimOv][kPrimOv];
            qbx_ov1d_t<kPrimOv>(PA[t], PB[t], inv2p, la[t], nj, S);
            const int ii = la[t], j = lb[t];
            if (g.kind == ONE_MULTIPOLE) {
                // (x - M)^d = sum_k C(d, k) (x - B)^k (B - M)^(d - k)
                const int d = g.deg[t];
                const double bm = B[t] - g.C[t];
                double m = 0.0, c = 1.0;
                for (int k = 0; k <= d; ++k) {
                    double pw = 1.0;
                    for (int q = 0; q < d - k; ++q) pw *= bm;
                    m += c * pw * S[ii][j + k];
                    c = c * (d - k) / (k + 1);
                }
                f[t] = m;
            } else if (g.kind == ONE_KINETIC) {
                // d^2/dx^2 acting on the second primitive
                double r = -2.0 * b * (2 * j + 1) * S[ii][j] + 4.0 * b * b * S[ii][j + 2];
                if (j > 1) r += (double)(j * (j - 1)) * S[ii][j - 2];
                f[t] = r;
                s0[t] = S[ii][j];
            } else {
                f[t] = S[ii][j];
            }
        }
        const double pre = 5.568327996831707845 * kab / (p * sqrt(p));   // (pi / p)^(3/2) e^{-ab/p r^2}
        if (g.kind == ONE_KINETIC)
            val = pre * -0.5 * (f[0] * s0[1] * s0[2] + s0[0] * f[1] * s0[2] + s0[0] * s0[1] * f[2]);
        else
            val = pre * f[0] * f[1] * f[2];
    }
    g.out[i] = val;
}

// ---------------------------------------------------------------------------
// kernel: one-electron part of the RHF energy gradient
// ---------------------------------------------------------------------------
struct OneGradArgs {
    const DevShell* sh;
    const PrimPair* pp;
    const int* ppk;
    const int* pA;
    const int* pB;
    const int* po;
    const int* pn;
    int npair;
    int N;
    const double* boys;
    const double* bexp;
    int nnuc;
    const double* Z;
    const double* R;
    const double* Dt;      // total density
    const double* W;       // energy-weighted density
    const int* poff;
    double* g_xpn;
    double* g_lin;
    double* g_cen;
};

constexpr int kGradPairPrims = QBX_GRAD_MAXP * QBX_GRAD_MAXP;   // primitive pairs per shell pair in the 1e gradient grid

// D (T + V) - W S for one primitive pair and Cartesian components la, lb of
// the two primitives (any angular momentum up to 2 above the shells').
__device__ double one_e_energy(const int* la, const int* lb, const PrimPair& P, const double* AB,
                               const OneGradArgs& a, double wd, double ww)
{
    const double PA[3] = {P.PAx, P.PAy, P.PAz};
    const double PB[3] = {P.PAx + AB[0], P.PAy + AB[1], P.PAz + AB[2]};
    double S[3][QBX_OV_MAX][QBX_OV_MAX];
    for (int t = 0; t < 3; ++t) qbx_ov1d(PA[t], PB[t], P.inv2p, la[t], lb[t] + 2, S[t]);
    const double sx = S[0][la[0]][lb[0]], sy = S[1][la[1]][lb[1]], sz = S[2][la[2]][lb[2]];
    const double b = P.pad;
    double d2[3];
    for (int t = 0; t < 3; ++t) {
        const int j = lb[t];
        double r = -2.0 * b * (2 * j + 1) * S[t][la[t]][j] + 4.0 * b * b * S[t][la[t]][j + 2];
        if (j > 1) r += (double)(j * (j - 1)) * S[t][la[t]][j - 2];
        d2[t] = r;
    }
    const double pre = 5.568327996831707845 * P.Kp * rsqrt(P.p);
    const double ovl = pre * sx * sy * sz;
    const double kin = pre * -0.5 * (d2[0] * sy * sz + sx * d2[1] * sz + sx * sy * d2[2]);
    double nuc = 0.0;
    if (wd != 0.0) {
        const int Lt = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2];
        for (int n = 0; n < a.nnuc; ++n) {
            const double PC[3] = {P.Px - a.R[3 * n], P.Py - a.R[3 * n + 1], P.Pz - a.R[3 * n + 2]};
            const double T = P.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
            double F[9];   // s and p shells (the gradient rejects higher L), raised by up to two orders
            qbx_boys_rt(Lt, T, F, a.boys, a.bexp);
            nuc += a.Z[n] * coulomb1_comp_t<8>(la, lb, PA, PC, AB, P.inv2p, F, Lt);
        }
        nuc *= -2.0 * kPi * P.Kp;
    }
    return wd * (kin + nuc) - ww * ovl;
}

// One thread per (shell pair, primitive pair).  With I(a, b) an integral over
// the primitives' Cartesian components a, b (weights included):
//   dI/dA_i = 2 alpha I(a + 1_i, b) - a_i I(a - 1_i, b),
//   dI/dalpha = -sum_i I(a + 2_i, b)   (d/dalpha exp(-alpha r_A^2) = -r_A^2 exp(...)),
// likewise for B, beta; and c dI/dc = I for either primitive.
__global__ __launch_bounds__(256) void one_e_grad_kernel(OneGradArgs a)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int pair = gid / kGradPairPrims, u = gid % kGradPairPrims;
    if (pair >= a.npair || u >= a.pn[pair]) return;
    const int sA = a.pA[pair], sB = a.pB[pair];
    const DevShell A = a.sh[sA];
    const DevShell B = a.sh[sB];
    const PrimPair P = a.pp[a.po[pair] + u];
    const int kab = a.ppk[a.po[pair] + u];
    const int ka = kab & 0xffff, kb = kab >> 16;
    const double alpha = P.p - P.pad, beta = P.pad;
    const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
    const double fac = sA == sB ? 1.0 : 2.0;
    const size_t N = (size_t)a.N;
    double gxa = 0.0, gxb = 0.0, gl = 0.0, gAc[3] = {0.0, 0.0, 0.0}, gBc[3] = {0.0, 0.0, 0.0};
    for (int ca = 0; ca < qbx_ncart(A.L); ++ca)
        for (int cb = 0; cb < qbx_ncart(B.L); ++cb) {
            const int ia = A.bf[ca], ib = B.bf[cb];
            if (ia < 0 || ib < 0) continue;
            int la[3], lb[3];
            qbx_cart(A.L, ca, la);
            qbx_cart(B.L, cb, lb);
            const double sc = fac * A.sc[ca] * B.sc[cb];
            const double wd = sc * a.Dt[ia * N + ib], ww = sc * a.W[ia * N + ib];
            gl += one_e_energy(la, lb, P, AB, a, wd, ww);
            for (int i = 0; i < 3; ++i) {
                int l2[3] = {la[0], la[1], la[2]};
                l2[i] += 2;
                gxa -= one_e_energy(l2, lb, P, AB, a, wd, ww);
                l2[i] -= 1;
                double g = 2.0 * alpha * one_e_energy(l2, lb, P, AB, a, wd, ww);
                if (la[i] > 0) {
                    l2[i] -= 2;
                    g -= la[i] * one_e_energy(l2, lb, P, AB, a, wd, ww);
                }
                gAc[i] += g;
                int m2[3] = {lb[0], lb[1], lb[2]};
                m2[i] += 2;
                gxb -= one_e_energy(la, m2, P, AB, a, wd, ww);
                m2[i] -= 1;
                g = 2.0 * beta * one_e_energy(la, m2, P, AB, a, wd, ww);
                if (lb[i] > 0) {
                    m2[i] -= 2;
                    g -= lb[i] * one_e_energy(la, m2, P, AB, a, wd, ww);
                }
                gBc[i] += g;
            }
        }
    const int pA = a.poff[sA], pB = a.poff[sB];
    atomicAdd(a.g_xpn + pA + ka, gxa);
    atomicAdd(a.g_xpn + pB + kb, gxb);
    atomicAdd(a.g_lin + pA + ka, gl);
    atomicAdd(a.g_lin + pB + kb, gl);
    for (int i = 0; i < 3; ++i) {
        atomicAdd(a.g_cen + 3 * sA + i, gAc[i]);
        atomicAdd(a.g_cen + 3 * sB + i, gBc[i]);
    }
}

// One thread per (shell pair, component pair); a.mc components per shell
// (those of the basis's highest L).  NM: twice the highest L (the nuclear
// attraction's recurrence size).
template <int KIND, int NM>
__global__ __launch_bounds__(256) void one_e_kernel(OneArgs a)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int pair = gid / (a.mc * a.mc);
    const int slot = gid % (a.mc * a.mc);
    if (pair >= a.npair) return;
    const DevShell A = a.sh[a.pA[pair]];
    const DevShell B = a.sh[a.pB[pair]];
    const int ca = slot / a.mc, cb = slot % a.mc;
    if (ca >= qbx_ncart(A.L) || cb >= qbx_ncart(B.L)) return;
    const int ia = A.bf[ca], ib = B.bf[cb];
    if (ia < 0 || ib < 0) return;
    int la[3], lb[3];
    qbx_cart(A.L, ca, la);
    qbx_cart(B.L, cb, lb);
    const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
    double acc = 0.0, accB = 0.0;
    for (int u = 0; u < a.pn[pair]; ++u) {
        const PrimPair P = a.pp[a.po[pair] + u];
        const double PA[3] = {P.PAx, P.PAy, P.PAz};
        const double PB[3] = {P.PAx + AB[0], P.PAy + AB[1], P.PAz + AB[2]};
        if (KIND == ONE_NUCLEAR_OWN) {
            // row function on A: A's own nucleus; row function on B: B's
            const int Lt = A.L + B.L;
            for (int side = 0; side < 2; ++side) {
                const int n = a.own[side ? a.pB[pair] : a.pA[pair]];
                if (n < 0) continue;
                const double PC[3] = {P.Px - a.R[3 * n], P.Py - a.R[3 * n + 1], P.Pz - a.R[3 * n + 2]};
                const double T = P.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                double F[NM + 1];
                qbx_boys_rt(Lt, T, F, a.boys, a.bexp);
                const double v = -2.0 * kPi * P.Kp * a.Z[n] * coulomb1_comp_t<NM>(la, lb, PA, PC, AB, P.inv2p, F, Lt);
                if (side) accB += v;
                else acc += v;
            }
        } else if (KIND == ONE_NUCLEAR) {
            const int Lt = A.L + B.L;
            const double Pc[3] = {P.Px, P.Py, P.Pz};
            double v = 0.0;
            for (int n = 0; n < a.nnuc; ++n) {
                const double PC[3] = {Pc[0] - a.R[3 * n], Pc[1] - a.R[3 * n + 1], Pc[2] - a.R[3 * n + 2]};
                const double T = P.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                double F[NM + 1];
                qbx_boys_rt(Lt, T, F, a.boys, a.bexp);
                v += a.Z[n] * coulomb1_comp_t<NM>(la, lb, PA, PC, AB, P.inv2p, F, Lt);
            }
            acc += -2.0 * kPi * P.Kp * v;
        } else {
            double S[3][QBX_OV_MAX][QBX_OV_MAX];
            const int extra = KIND == ONE_KINETIC ? 2 : (KIND == ONE_MULTIPOLE ? 0 : 0);
            double v;
            if (KIND == ONE_MULTIPOLE) {
                for (int t = 0; t < 3; ++t) qbx_ov1d(PA[t], PB[t], P.inv2p, la[t], lb[t] + a.deg[t], S[t]);
                v = 1.0;
                const double BM[3] = {B.x - a.M[0], B.y - a.M[1], B.z - a.M[2]};
                for (int t = 0; t < 3; ++t) {
                    const int n = a.deg[t];
                    double s = 0.0, binom = 1.0;
                    for (int k = 0; k <= n; ++k) {
                        double pw = 1.0;
                        for (int q = 0; q < n - k; ++q) pw *= BM[t];
                        s += binom * pw * S[t][la[t]][lb[t] + k];
                        binom = binom * (n - k) / (k + 1);
                    }
                    v *= s;
                }
            } else {
                for (int t = 0; t < 3; ++t) qbx_ov1d(PA[t], PB[t], P.inv2p, la[t], lb[t] + extra, S[t]);
                const double sx = S[0][la[0]][lb[0]], sy = S[1][la[1]][lb[1]], sz = S[2][la[2]][lb[2]];
                if (KIND == ONE_OVERLAP) {
                    v = sx * sy * sz;
                } else {
                    const double b = P.pad;
                    double d2[3];
                    for (int t = 0; t < 3; ++t) {
                        const int j = lb[t];
                        double r = -2.0 * b * (2 * j + 1) * S[t][la[t]][j] + 4.0 * b * b * S[t][la[t]][j + 2];
                        if (j > 1) r += (double)(j * (j - 1)) * S[t][la[t]][j - 2];
                        d2[t] = r;
                    }
                    v = -0.5 * (d2[0] * sy * sz + sx * d2[1] * sz + sx * sy * d2[2]);
                }
            }
            // (pi/p)^{3/2} c_a c_b K_AB = pi^{3/2} Kp / sqrt(p)
            acc += 5.568327996831707845 * P.Kp * rsqrt(P.p) * v;
        }
    }
    acc *= A.sc[ca] * B.sc[cb];
    a.out[(size_t)ia * a.N + ib] = acc;
    a.out[(size_t)ib * a.N + ia] = KIND == ONE_NUCLEAR_OWN ? accB * (A.sc[ca] * B.sc[cb]) : acc;
}

// max over the density blocks of shell pair (s1, s2): |D[s][i][j]|, all 2*nset
// matrices (DJ and DK sets), i in s1, j in s2
__global__ void pair_dmax_kernel(const DevShell* sh, int ns, const double* D, int nmat, int N, double* out)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ns * ns) return;
    const int s1 = g / ns, s2 = g % ns;
    const DevShell A = sh[s1], B = sh[s2];
    double m = 0.0;
    for (int ca = 0; ca < qbx_ncart(A.L); ++ca) {
        if (A.bf[ca] < 0) continue;
        for (int cb = 0; cb < qbx_ncart(B.L); ++cb) {
            if (B.bf[cb] < 0) continue;
            for (int k = 0; k < nmat; ++k)
                m = fmax(m, fabs(D[(size_t)k * N * N + (size_t)A.bf[ca] * N + B.bf[cb]]));
        }
    }
    out[g] = m;
}

__global__ void fill_zero_kernel(double* p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// G[s] = (J + J^T)/2 - (K + K^T)/2
// G = (J + J^T)/2 - (K + K^T)/2, J and K summed over their `copies` copies
// (stride NN * nset doubles).
__global__ void symmetrize_kernel(const double* J, const double* K, double* G, int N, int nset, int copies)
{
    const size_t NN = (size_t)N * N;
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NN * nset) return;
    const size_t s = g / NN, r = g % NN;
    const size_t i = r / N, j = r % N;
    double v = 0.0;
    for (int c = 0; c < copies; ++c) {
        const double* Js = J + (size_t)c * NN * nset + s * NN;
        const double* Ks = K + (size_t)c * NN * nset + s * NN;
        v += 0.5 * (Js[i * N + j] + Js[j * N + i]) - 0.5 * (Ks[i * N + j] + Ks[j * N + i]);
    }
    G[g] = v;
}

// Which XCD runs each block (one write per block).
__global__ void xcc_probe_kernel(int* out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = qbx_xcc_id();
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
void zero_async(double* p, size_t n, hipStream_t s)
{
    if (n == 0) return;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    QBX_HIP(hipGetLastError());
}

int pair_class_index(int la, int lb)
{
    for (int c = 0; c < QBX_N_PAIR_CLASSES; ++c)
        if (qbx_pair_class_l[c][0] == la && qbx_pair_class_l[c][1] == lb) return c;
    return -1;
}

int class_pc(int cls, int side)
{
    return pair_class_index(qbx_class_l[cls][2 * side], qbx_class_l[cls][2 * side + 1]);
}

bool X_is_Y(int cls) { return class_pc(cls, 0) == class_pc(cls, 1); }

PairClassDev pcdev(const PairClass& c)
{
    PairClassDev d;
    d.n = c.n();
    d.shA = c.d_shA.p;
    d.shB = c.d_shB.p;
    d.ppo = c.d_ppo.p;
    d.ppn = c.d_ppn.p;
    d.AB = c.d_AB.p;
    d.Q = c.d_Q.p;
    return d;
}

#define QBX_TAB_ENTRY(R, MODE)
#define QBX_E64JK(a, b, c, d) &launch_eri<a, b, c, d, double, MODE_JK>,
#define QBX_E32JK(a, b, c, d) &launch_eri<a, b, c, d, float, MODE_JK>,
#define QBX_EDENSE(a, b, c, d) &launch_eri<a, b, c, d, double, MODE_DENSE>,
const EriLaunch k_jk64[QBX_N_CLASSES] = {QBX_FOR_EACH_CLASS(QBX_E64JK)};
const EriLaunch k_jk32[QBX_N_CLASSES] = {QBX_FOR_EACH_CLASS(QBX_E32JK)};
const EriLaunch k_dense[QBX_N_CLASSES] = {QBX_FOR_EACH_CLASS(QBX_EDENSE)};

EriLaunch grad_launcher(int cls)
{
    const int* L = qbx_class_l[cls];
#define QBX_GR(a, b, c, d) \
    if (L[0] == a && L[1] == b && L[2] == c && L[3] == d) return &launch_eri_grad<a, b, c, d>;
    QBX_GR(0, 0, 0, 0) QBX_GR(1, 0, 0, 0) QBX_GR(1, 0, 1, 0) QBX_GR(1, 1, 0, 0) QBX_GR(1, 1, 1, 0) QBX_GR(1, 1, 1, 1)
#undef QBX_GR
    return nullptr;
}

EriLaunch schwarz_launcher(int la, int lb)
{
#define QBX_SW(a, b) if (la == a && lb == b) return &launch_eri<a, b, a, b, double, MODE_SCHWARZ>;
    QBX_SW(0, 0) QBX_SW(1, 0) QBX_SW(1, 1) QBX_SW(2, 0) QBX_SW(2, 1) QBX_SW(2, 2)
#undef QBX_SW
    return nullptr;
}

// The pair lists an ERI class runs over: the staged classes (few long waves
// with a large fixed cost each, where the extra partial waves of the buckets
// cost more than the even lanes save) take the plain Schwarz order of a
// bucketed ket class, and of the bra class when it is the same pair class.
const PairClass& pair_view(const qbx_basis* b, int cls, int side)
{
    const PairClass& P = b->pc[class_pc(cls, side)];
    if (P.plain && qbx_class_ws[cls] > 0 && (side == 1 || X_is_Y(cls))) return *P.plain;
    return P;
}

EriArgs base_args(qbx_basis* b, int cls)
{
    EriArgs a;
    std::memset(&a, 0, sizeof(a));
    const int X = class_pc(cls, 0), Y = class_pc(cls, 1);
    a.bra = pcdev(pair_view(b, cls, 0));
    a.ket = pcdev(pair_view(b, cls, 1));
    a.pp = b->d_pp.p;
    a.sh = b->d_sh.p;
    a.boys = b->ctx->boys.p;
    a.bexp = b->ctx->bexp.p;
    const int lt = qbx_class_l[cls][0] + qbx_class_l[cls][1] + qbx_class_l[cls][2] + qbx_class_l[cls][3];
    a.boysl = b->ctx->boysl.p + (size_t)lt * QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE;
    a.same = X == Y;
    a.N = b->nbf;
    a.nbk = pair_view(b, cls, 1).nbk;
    return a;
}

// Workspace of a staged class for up to `nwave` waves (kStageSlots at most).
void attach_workspace(qbx_basis* b, int cls, int nwave, EriArgs& a)
{
    const int wsv = qbx_class_ws[cls];
    if (wsv == 0 || nwave <= 0) return;
    int slots = std::min(nwave, kStageSlots);
    slots = (slots + 3) / 4 * 4;
    b->ws[cls].ensure((size_t)slots * wsv * 64);
    a.ws = b->ws[cls].p;
    a.ws_slots = (int)(b->ws[cls].n / ((size_t)wsv * 64));
}

// Schwarz bounds, then every pair class sorted by descending bound.
// Twin bras.  Shells that share centre, L and exponents form a family (the
// segmented copies of a general contraction); shell pairs of one pair class
// whose shells come from the same two families and that kept the same
// primitive pairs have identical primitive geometry and differ only in the
// c_a c_b factors.  Groups of up to TWIN_MAX such pairs, in Schwarz order.
void find_twins(qbx_basis* b)
{
    const int ns = (int)b->shells.size();
    std::map<std::vector<double>, int> fam_of;
    std::vector<int> fam(ns);
    for (int s = 0; s < ns; ++s) {
        const HostShell& H = b->shells[s];
        std::vector<double> key = {H.c[0], H.c[1], H.c[2], (double)H.L, (double)H.nprim};
        for (int k = 0; k < H.nprim; ++k) key.push_back(b->pexp[H.poff + k]);
        auto it = fam_of.find(key);
        if (it == fam_of.end()) it = fam_of.emplace(key, (int)fam_of.size()).first;
        fam[s] = it->second;
    }
    for (int c = 0; c < QBX_N_PAIR_CLASSES; ++c) {
        const PairClass& P = b->pc[c];
        std::map<std::vector<int>, std::vector<int>> sig;
        for (int k = 0; k < P.n(); ++k) {
            std::vector<int> key = {fam[P.shA[k]], fam[P.shB[k]], P.ppn[k]};
            for (int u = 0; u < P.ppn[k]; ++u) key.push_back(b->ppk[P.ppo[k] + u]);
            sig[key].push_back(k);
        }
        b->twins[c].clear();
        for (auto& kv : sig) {
            const std::vector<int>& m = kv.second;
            for (size_t o = 0; o + 1 < m.size(); o += TWIN_MAX) {
                const size_t e = std::min(m.size(), o + (size_t)TWIN_MAX);
                if (e - o >= 2) b->twins[c].emplace_back(m.begin() + o, m.begin() + e);
            }
        }
        std::sort(b->twins[c].begin(), b->twins[c].end(),
                  [](const std::vector<int>& x, const std::vector<int>& y) { return x[0] < y[0]; });
    }
}

bool high_class(int c) { return qbx_pair_class_l[c][0] > QBX_GEN_LMAX; }

GenArgs gen_base(const qbx_basis* b)
{
    GenArgs g;
    std::memset(&g, 0, sizeof(g));
    g.sh = b->d_sh.p;
    g.pp = b->d_pp.p;
    g.ppk = b->d_ppk.p;
    g.poff = b->d_poff.p;
    g.pexp = b->d_pexp.p;
    g.boys = b->ctx->boys.p;
    g.bexp = b->ctx->bexp.p;
    g.N = b->nbf;
    return g;
}

// Schwarz bounds of the pairs of a pair class above the generated classes:
// max |(ab|ab)| over the pair's components from the generic kernel, in the
// class's current order.
std::vector<double> high_class_schwarz(qbx_basis* b, int c, hipStream_t st)
{
    PairClass& P = b->pc[c];
    const int n = P.n();
    DevBuf<int64_t> pre;
    DevBuf<unsigned long long> bits;
    std::vector<int64_t> h(n + 1);
    std::iota(h.begin(), h.end(), (int64_t)0);
    pre.upload(h, st);
    bits.ensure(n);
    QBX_HIP(hipMemsetAsync(bits.p, 0, n * sizeof(unsigned long long), st));
    GenArgs g = gen_base(b);
    g.nbra = n;
    g.total = n;
    g.prefix = pre.p;
    g.hA = P.d_shA.p;
    g.hB = P.d_shB.p;
    g.ho = P.d_ppo.p;
    g.hn = P.d_ppn.p;
    g.qbits = bits.p;
    launch_eri_generic(g, 2 * (qbx_pair_class_l[c][0] + qbx_pair_class_l[c][1]), MODE_SCHWARZ, st);
    std::vector<unsigned long long> hb(n);
    QBX_HIP(hipMemcpyAsync(hb.data(), bits.p, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    std::vector<double> Q(n);
    for (int i = 0; i < n; ++i) {
        double v;
        std::memcpy(&v, &hb[i], sizeof(double));
        Q[i] = std::sqrt(v);
    }
    return Q;
}

void ensure_schwarz(qbx_basis* b)
{
    if (b->schwarz) return;
    hipStream_t st = b->ctx->stream;
    std::vector<DevBuf<double>> q(QBX_N_PAIR_CLASSES);
    std::vector<std::vector<double>> qhigh(QBX_N_PAIR_CLASSES);
    for (int c = 0; c < QBX_N_PAIR_CLASSES; ++c) {
        PairClass& P = b->pc[c];
        if (P.n() == 0) continue;
        if (high_class(c)) {
            qhigh[c] = high_class_schwarz(b, c, st);
            continue;
        }
        q[c].ensure(P.n());
        EriArgs a;
        std::memset(&a, 0, sizeof(a));
        a.bra = pcdev(P);
        a.ket = pcdev(P);
        a.pp = b->d_pp.p;
        a.sh = b->d_sh.p;
        a.boys = b->ctx->boys.p;
        a.bexp = b->ctx->bexp.p;
        const int lt = 2 * (qbx_pair_class_l[c][0] + qbx_pair_class_l[c][1]);
        a.boysl = b->ctx->boysl.p + (size_t)lt * QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE;
        a.total = P.n();
        a.nbra = P.n();
        a.nbk = 1;
        a.same = 1;
        a.N = b->nbf;
        a.qout = q[c].p;
        EriLaunch f = schwarz_launcher(qbx_pair_class_l[c][0], qbx_pair_class_l[c][1]);
        f(a, st);
    }
    for (int c = 0; c < QBX_N_PAIR_CLASSES; ++c) {
        PairClass& P = b->pc[c];
        const int n = P.n();
        if (n == 0) {
            P.nbk = 1;
            P.bstart.assign(2, 0);
            continue;
        }
        std::vector<double> Q(n);
        if (high_class(c)) {
            Q = qhigh[c];
        } else {
            QBX_HIP(hipMemcpyAsync(Q.data(), q[c].p, n * sizeof(double), hipMemcpyDeviceToHost, st));
            QBX_HIP(hipStreamSynchronize(st));
        }
        std::vector<int> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return Q[x] > Q[y]; });
        std::vector<int> qpos(n);
        for (int k = 0; k < n; ++k) qpos[perm[k]] = k;
        // buckets where they make the lanes of a wave markedly more even:
        // busy fraction of the primitive loop over chunks of 64 pairs, in
        // Schwarz order and within buckets.  Measured: 24 H2O in cc-pVDZ
        // gains 1.9-2.8x on this estimate for its s/p pair classes and its
        // Fock build drops 80 -> 54 ms; the 592-function cluster of
        // 3-primitive shells gains at most 1.29x, and bucketing its (pp)
        // kets at that ratio made (pp|pp) 1.55 -> 2.17 ms (extra partial
        // waves per bra), so the bar is 1.5x.
        auto busy = [&](const std::vector<int>& order, bool by_bucket) {
            double work = 0.0, span = 0.0;
            int mx = 0, fill = 0, prev = -1;
            for (size_t k = 0; k < order.size(); ++k) {
                const int pn = P.ppn[order[k]];
                const int bk = by_bucket ? ket_bucket(pn) : 0;
                if (fill == 64 || (fill > 0 && bk != prev)) {   // chunk full, or a bucket starts its own
                    span += 64.0 * mx;
                    mx = 0;
                    fill = 0;
                }
                prev = bk;
                work += pn;
                mx = std::max(mx, pn);
                ++fill;
            }
            span += 64.0 * mx;
            return span > 0.0 ? work / span : 1.0;
        };
        const double busy_plain = busy(perm, false);
        std::vector<int> bperm = perm;
        std::stable_sort(bperm.begin(), bperm.end(),
                         [&](int x, int y) { return ket_bucket(P.ppn[x]) < ket_bucket(P.ppn[y]); });
        const bool bucketed = b->ctx->ket_buckets && busy(bperm, true) >= 1.5 * busy_plain;
        if (bucketed) {
            std::unique_ptr<PairClass> pl(new PairClass());
            for (int k = 0; k < n; ++k) {
                const int o = perm[k];
                pl->shA.push_back(P.shA[o]);
                pl->shB.push_back(P.shB[o]);
                pl->ppo.push_back(P.ppo[o]);
                pl->ppn.push_back(P.ppn[o]);
                pl->Q.push_back(Q[o]);
                for (int t = 0; t < 3; ++t) pl->AB.push_back(P.AB[3 * o + t]);
                pl->rank.push_back(k);
            }
            pl->bstart = {0, n};
            pl->d_shA.upload(pl->shA, st);
            pl->d_shB.upload(pl->shB, st);
            pl->d_ppo.upload(pl->ppo, st);
            pl->d_ppn.upload(pl->ppn, st);
            pl->d_AB.upload(pl->AB, st);
            pl->d_Q.upload(pl->Q, st);
            P.plain = std::move(pl);
            perm.swap(bperm);
        }
        PairClass S;
        S.nbk = bucketed ? kKetBuckets : 1;
        S.bstart.assign(S.nbk + 1, 0);
        for (int k = 0; k < n; ++k) {
            const int o = perm[k];
            S.rank.push_back(qpos[o]);
            ++S.bstart[(bucketed ? ket_bucket(P.ppn[o]) : 0) + 1];
            S.shA.push_back(P.shA[o]);
            S.shB.push_back(P.shB[o]);
            S.ppo.push_back(P.ppo[o]);
            S.ppn.push_back(P.ppn[o]);
            S.Q.push_back(Q[o]);
            for (int t = 0; t < 3; ++t) S.AB.push_back(P.AB[3 * o + t]);
        }
        P.shA.swap(S.shA);
        P.shB.swap(S.shB);
        P.ppo.swap(S.ppo);
        P.ppn.swap(S.ppn);
        P.Q.swap(S.Q);
        P.AB.swap(S.AB);
        for (int t = 0; t < S.nbk; ++t) S.bstart[t + 1] += S.bstart[t];
        P.rank.swap(S.rank);
        P.bstart.swap(S.bstart);
        P.nbk = S.nbk;
        P.d_shA.upload(P.shA, st);
        P.d_shB.upload(P.shB, st);
        P.d_ppo.upload(P.ppo, st);
        P.d_ppn.upload(P.ppn, st);
        P.d_AB.upload(P.AB, st);
        P.d_Q.upload(P.Q, st);
    }
    QBX_HIP(hipStreamSynchronize(st));
    b->schwarz = true;
    find_twins(b);
}

// number of leading entries of a descending array that are >= v
int count_ge(const std::vector<double>& q, double v)
{
    return (int)(std::upper_bound(q.begin(), q.end(), v, [](double x, double y) { return x > y; }) - q.begin());
}

// Screened task lists: for bra i of class X, kets j of class Y with
// Q_i Q_j >= thr (and j before or at i in Schwarz order when X == Y).  Under
// mixed precision the surviving kets split at Q_i Q_j >= cut into an FP64 head
// and an FP32 tail.  One task range per (bra, ket bucket): task t = i nbk + k.
struct TaskCounts {
    std::vector<int64_t> p64, p32;
    std::vector<int> k64, k32;
    std::vector<int> w64, w32;   // waves of 64 kets per task range, prefix sums
    std::vector<int> s64, s32;   // K-row blocks per task range, prefix sums
    int ntb = 0;                 // task ranges (bras x buckets)
    int rc = ROW_CHUNKS;         // chunks per K-row block
};

// Surviving kets of bra i in ket bucket k: FP64 [0, c64), FP32 [c64, cnt),
// relative to the bucket's first position.
void bra_range(const qbx_basis* b, int cls, int i, int k, double thr, int mode, double cut, int& c64, int& cnt)
{
    const int X = class_pc(cls, 0), Y = class_pc(cls, 1);
    const PairClass& PX = pair_view(b, cls, 0);
    const PairClass& PY = pair_view(b, cls, 1);
    const int s0 = PY.bstart[k], s1 = PY.bstart[k + 1];
    const double qi = PX.Q[i];
    auto ge = [&](double v) {
        return (int)(std::upper_bound(PY.Q.begin() + s0, PY.Q.begin() + s1, v,
                                      [](double x, double y) { return x > y; }) - (PY.Q.begin() + s0));
    };
    if (thr <= 0.0) cnt = s1 - s0;
    else cnt = qi > 0.0 ? ge(thr / qi) : 0;
    if (X == Y)
        cnt = std::min(cnt, (int)(std::upper_bound(PY.rank.begin() + s0, PY.rank.begin() + s1, PX.rank[i]) -
                                  (PY.rank.begin() + s0)));
    c64 = cnt;
    if (mode == QBX_PREC_MIXED) c64 = std::min(cnt, qi > 0.0 ? ge(cut / qi) : 0);
}

// chunks of 64 kets per K-row task of class cls
int task_chunks(int cls)
{
    if (qbx_class_ws[cls] == 0) return ROW_CHUNKS;
    return qbx_class_split[cls] > 1 && qbx_class_split_chunks[cls] > 0 ? qbx_class_split_chunks[cls]
                                                                       : STAGED_ROW_CHUNKS;
}

// start64: per task range, the first FP64 ket left to these lists (the twin
// groups take the kets before it); null: all of them
void task_counts(const qbx_basis* b, int cls, double thr, int mode, double cut, TaskCounts& tc,
                 const std::vector<int>* start64 = nullptr)
{
    const int nx = b->pc[class_pc(cls, 0)].n();
    const PairClass& PY = pair_view(b, cls, 1);
    const int nk = PY.nbk;
    const int nt = nx * nk;
    const int rc = task_chunks(cls);   // chunks per K-row task
    tc.ntb = nt;
    tc.rc = rc;
    tc.p64.assign(nt + 1, 0);
    tc.p32.assign(nt + 1, 0);
    tc.k64.assign(std::max(nt, 1), 0);
    tc.k32.assign(std::max(nt, 1), 0);
    tc.w64.assign(nt + 1, 0);
    tc.w32.assign(nt + 1, 0);
    tc.s64.assign(nt + 1, 0);
    tc.s32.assign(nt + 1, 0);
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nk; ++k) {
            const int t = i * nk + k;
            int c64, cnt;
            bra_range(b, cls, i, k, thr, mode, cut, c64, cnt);
            const int k0 = start64 ? std::min((*start64)[t], c64) : 0;
            const int s0 = PY.bstart[k];
            tc.k64[t] = s0 + k0;
            tc.k32[t] = s0 + c64;
            tc.p64[t + 1] = tc.p64[t] + (c64 - k0);
            tc.p32[t + 1] = tc.p32[t] + (cnt - c64);
            tc.w64[t + 1] = tc.w64[t] + (c64 - k0 + 63) / 64;
            tc.w32[t + 1] = tc.w32[t] + (cnt - c64 + 63) / 64;
            tc.s64[t + 1] = tc.s64[t] + ((c64 - k0 + 63) / 64 + rc - 1) / rc;
            tc.s32[t + 1] = tc.s32[t] + ((cnt - c64 + 63) / 64 + rc - 1) / rc;
        }
}

void upload_tasks(TaskSet& t64, TaskSet& t32, const TaskCounts& tc, hipStream_t st)
{
    const int nt = tc.ntb;
    t64.prefix.upload(tc.p64, st);
    t64.koff.upload(tc.k64, st);
    t64.cpre.upload(tc.w64, st);
    t64.spre.upload(tc.s64, st);
    t64.nsuper = tc.s64[nt];
    t64.chunks = tc.rc;
    t64.total = tc.p64[nt];
    t64.nwave = tc.w64[nt];
    t64.nbra = nt;
    t32.prefix.upload(tc.p32, st);
    t32.koff.upload(tc.k32, st);
    t32.cpre.upload(tc.w32, st);
    t32.spre.upload(tc.s32, st);
    t32.nsuper = tc.s32[nt];
    t32.chunks = tc.rc;
    t32.total = tc.p32[nt];
    t32.nwave = tc.w32[nt];
    t32.nbra = nt;
}

// Task lists of the twin groups of class cls: the FP64 kets that every member
// of a group keeps, [0, common), in chunks of 64, ROW_CHUNKS chunks a task.
// Each member's kets beyond the common prefix, and all FP32 kets, stay with
// the per-bra lists (start64 receives the common prefix of every member), so
// the shared evaluation never covers a ket that a member would not.
void twin_tasks(const qbx_basis* b, int cls, double thr, int mode, double cut, TwinSet& w64,
                std::vector<int>& start64, hipStream_t st)
{
    const std::vector<std::vector<int>>& groups = b->twins[class_pc(cls, 0)];
    const PairClass& PY = pair_view(b, cls, 1);
    std::vector<int> pre(1, 0), mem;
    int twmax = 0;
    for (const std::vector<int>& m : groups)
        for (int k = 0; k < PY.nbk; ++k) {
            int common = INT32_MAX;
            for (int i : m) {
                int c64, cnt;
                bra_range(b, cls, i, k, thr, mode, cut, c64, cnt);
                common = std::min(common, c64);
            }
            if (common <= 0) continue;
            twmax = std::max(twmax, (int)m.size());
            const int s0 = PY.bstart[k];
            for (size_t t = 0; t < (size_t)TWIN_MAX; ++t) {
                const bool in = t < m.size();
                mem.push_back(in ? m[t] : -1);
                mem.push_back(in ? s0 : 0);
                mem.push_back(in ? s0 + common : 0);
                if (in) start64[m[t] * PY.nbk + k] = common;
            }
            const int ch = (common + 63) / 64;
            pre.push_back(pre.back() + (ch + ROW_CHUNKS - 1) / ROW_CHUNKS);
        }
    const int ng = (int)pre.size() - 1;
    w64.gpre.upload(pre, st);
    if (ng > 0) w64.gmem.upload(mem, st);
    w64.ngroup = ng;
    w64.ntask = pre.back();
    w64.twmax = twmax;
}

// Estimated work of a class's task lists: primitive shell quartets and
// contracted shell quartets weighted by the class's cost per unit.
double class_cost(const qbx_basis* b, int cls, const TaskCounts& tc)
{
    const PairClass& PX = pair_view(b, cls, 0);
    const PairClass& PY = pair_view(b, cls, 1);
    std::vector<int64_t> kp(PY.n() + 1, 0);
    for (int j = 0; j < PY.n(); ++j) kp[j + 1] = kp[j] + PY.ppn[j];
    double prim = 0.0, quart = 0.0;
    for (int t = 0; t < tc.ntb; ++t) {
        const int n64 = (int)(tc.p64[t + 1] - tc.p64[t]), n32 = (int)(tc.p32[t + 1] - tc.p32[t]);
        const int64_t kets = (kp[tc.k64[t] + n64] - kp[tc.k64[t]]) + (kp[tc.k32[t] + n32] - kp[tc.k32[t]]);
        prim += (double)PX.ppn[t / PY.nbk] * (double)kets;
        quart += n64 + n32;
    }
    return prim * qbx_class_cost_prim[cls] + quart * qbx_class_cost_quartet[cls];
}

void ensure_tasks(qbx_basis* b, double thr, int mode, double cut)
{
    if (mode != QBX_PREC_MIXED) cut = 0.0;
    if (b->tasks_valid && b->task_thr == thr && b->task_mode == mode && b->task_cut == cut) return;
    // the lists are rewritten (and may be reallocated) on the context stream:
    // the last build, on whatever stream it ran, must be done reading them
    if (b->built) QBX_HIP(hipEventSynchronize(b->ev_done));
    hipStream_t st = b->ctx->stream;
    for (int cls = 0; cls < QBX_N_CLASSES; ++cls) {
        TaskCounts tc;
        task_counts(b, cls, thr, mode, cut, tc);
        const int X = class_pc(cls, 0);
        const int nx = b->pc[X].n();
        upload_tasks(b->t64[cls], b->t32[cls], tc, st);
        b->class_cost[cls] = class_cost(b, cls, tc);
        b->use_twin[cls] = false;
        // only where bra and ket classes differ: within one pair class a bra's
        // kets end at its own Schwarz-sorted position, so the members' common
        // prefix is short and the group's fixed cost per block outweighs it
        // (measured on 24 H2O in cc-pVDZ: (ps|ss) 27.8 -> 24.7 ms, while
        // (ss|ss) 15.2 -> 16.0 and (ps|ps) 11.3 -> 12.0 ms)
        if (b->ctx->twins && qbx_class_twin[cls] && X != class_pc(cls, 1) && !b->twins[X].empty()) {
            std::vector<int> start64((size_t)nx * pair_view(b, cls, 1).nbk, 0);
            twin_tasks(b, cls, thr, mode, cut, b->w64[cls], start64, st);
            if (b->w64[cls].ngroup > 0) {
                TaskCounts ts;
                task_counts(b, cls, thr, mode, cut, ts, &start64);
                upload_tasks(b->s64[cls], b->s32[cls], ts, st);
                b->use_twin[cls] = true;
            }
        }
    }
    QBX_HIP(hipStreamSynchronize(st));
    b->tasks_valid = true;
    b->task_thr = thr;
    b->task_mode = mode;
    b->task_cut = cut;
}

// Quartet lists of the generic kernel at threshold thr: bras are the pairs of
// the classes above the generated ones in descending Schwarz order; bra i
// takes every lower pair with Q_i Q_j >= thr, then the higher pairs j <= i
// with Q_i Q_j >= thr (each higher-higher quartet once).
void ensure_hl_tasks(qbx_basis* b, double thr)
{
    if (!b->has_high) return;
    auto& H = b->hl;
    if (H.valid && H.thr == thr) return;
    if (b->built) QBX_HIP(hipEventSynchronize(b->ev_done));
    struct Pr {
        int A, B, o, n;
        double Q;
    };
    std::vector<Pr> hi, lo;
    for (int c = 0; c < QBX_N_PAIR_CLASSES; ++c) {
        const PairClass& P = b->pc[c];
        auto& dst = high_class(c) ? hi : lo;
        for (int i = 0; i < P.n(); ++i) dst.push_back({P.shA[i], P.shB[i], P.ppo[i], P.ppn[i], P.Q[i]});
    }
    auto byQ = [](const Pr& x, const Pr& y) { return x.Q > y.Q; };
    std::stable_sort(hi.begin(), hi.end(), byQ);
    std::stable_sort(lo.begin(), lo.end(), byQ);
    std::vector<double> hq(hi.size()), lq(lo.size());
    for (size_t i = 0; i < hi.size(); ++i) hq[i] = hi[i].Q;
    for (size_t i = 0; i < lo.size(); ++i) lq[i] = lo[i].Q;
    const int nb = (int)hi.size();
    std::vector<int64_t> pre(nb + 1, 0);
    std::vector<int> nlow(nb);
    for (int i = 0; i < nb; ++i) {
        int nl, nh;
        if (thr <= 0.0) {
            nl = (int)lo.size();
            nh = i + 1;
        } else if (hq[i] > 0.0) {
            nl = count_ge(lq, thr / hq[i]);
            nh = std::min(i + 1, count_ge(hq, thr / hq[i]));
        } else {
            nl = nh = 0;
        }
        nlow[i] = nl;
        pre[i + 1] = pre[i] + nl + nh;
    }
    hipStream_t st = b->ctx->stream;
    auto up = [&](const std::vector<Pr>& v, DevBuf<int>& A, DevBuf<int>& B, DevBuf<int>& o, DevBuf<int>& n,
                  DevBuf<double>& Q, const std::vector<double>& q) {
        std::vector<int> a(v.size()), bb(v.size()), oo(v.size()), nn(v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            a[i] = v[i].A;
            bb[i] = v[i].B;
            oo[i] = v[i].o;
            nn[i] = v[i].n;
        }
        A.upload(a, st);
        B.upload(bb, st);
        o.upload(oo, st);
        n.upload(nn, st);
        Q.upload(q, st);
    };
    up(hi, H.hA, H.hB, H.ho, H.hn, H.hQ, hq);
    up(lo, H.lA, H.lB, H.lo, H.ln, H.lQ, lq);
    H.nlow.upload(nlow, st);
    H.prefix.upload(pre, st);
    QBX_HIP(hipStreamSynchronize(st));
    H.nbra = nb;
    H.total = pre[nb];
    H.thr = thr;
    H.valid = true;
}

GenArgs gen_lists(const qbx_basis* b)
{
    GenArgs g = gen_base(b);
    const auto& H = b->hl;
    g.nbra = H.nbra;
    g.total = H.total;
    g.prefix = H.prefix.p;
    g.nlow = H.nlow.p;
    g.hA = H.hA.p;
    g.hB = H.hB.p;
    g.ho = H.ho.p;
    g.hn = H.hn.p;
    g.hQ = H.hQ.p;
    g.lA = H.lA.p;
    g.lB = H.lB.p;
    g.lo = H.lo.p;
    g.ln = H.ln.p;
    g.lQ = H.lQ.p;
    return g;
}

// Before a Fock build on stream st touches the basis's work buffers: a
// previous build still running on another stream must finish first (stream
// order covers a build on the same stream).
void order_after_last_build(qbx_basis* b, hipStream_t st)
{
    if (b->built && b->last_stream != st) QBX_HIP(hipEventSynchronize(b->ev_done));
}

void check_ctx_basis(qbx_ctx* ctx, const qbx_basis* b)
{
    if (!ctx) throw QbxError(QBX_EINVAL, "null context");
    if (!b) throw QbxError(QBX_EINVAL, "null basis");
    if (b->ctx != ctx) throw QbxError(QBX_EINVAL, "basis belongs to another context");
}

void run_one_e(qbx_ctx* ctx, const qbx_basis* b, int kind, int nnuc, const double* Z, const double* R,
               const double* M, const int32_t* deg, double* out)
{
    check_ctx_basis(ctx, b);
    if (!out) throw QbxError(QBX_EINVAL, "null output");
    QBX_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t NN = (size_t)b->nbf * b->nbf;
    DevBuf<double> dout, dZ, dR;
    dout.ensure(NN);
    zero_async(dout.p, NN, st);
    OneArgs a;
    std::memset(&a, 0, sizeof(a));
    a.sh = b->d_sh.p;
    a.pp = b->d_pp.p;
    a.pA = b->d_p1A.p;
    a.pB = b->d_p1B.p;
    a.po = b->d_p1o.p;
    a.pn = b->d_p1n.p;
    a.npair = (int)b->p1.size();
    a.N = b->nbf;
    a.boys = ctx->boys.p;
    a.bexp = ctx->bexp.p;
    a.out = dout.p;
    DevBuf<int> down;
    if (kind == ONE_NUCLEAR || kind == ONE_NUCLEAR_OWN) {
        if (nnuc <= 0 || !Z || !R) throw QbxError(QBX_EINVAL, "nuclear attraction needs at least one nucleus");
        std::vector<double> z(Z, Z + nnuc), r(R, R + 3 * nnuc);
        dZ.upload(z, st);
        dR.upload(r, st);
        a.nnuc = nnuc;
        a.Z = dZ.p;
        a.R = dR.p;
    }
    if (kind == ONE_NUCLEAR_OWN) {
        std::vector<int> own(b->shells.size(), -1);
        for (size_t s = 0; s < b->shells.size(); ++s)
            for (int n = 0; n < nnuc; ++n)
                if (std::fabs(b->shells[s].c[0] - R[3 * n]) + std::fabs(b->shells[s].c[1] - R[3 * n + 1]) +
                        std::fabs(b->shells[s].c[2] - R[3 * n + 2]) <= 1e-10) {
                    own[s] = n;
                    break;
                }
        down.upload(own, st);
        a.own = down.p;
    }
    if (kind == ONE_MULTIPOLE) {
        for (int t = 0; t < 3; ++t) {
            if (deg[t] < 0 || deg[t] > 4) throw QbxError(QBX_EUNSUP, "multipole degree per axis must be in [0, 4]");
            a.M[t] = M[t];
            a.deg[t] = deg[t];
        }
    }
    a.mc = qbx_ncart(b->maxL);
    const int64_t threads = (int64_t)a.npair * a.mc * a.mc;
    if (threads > 0) {
        if ((threads + 255) / 256 > 0x7fffffffLL) throw QbxError(QBX_EUNSUP, "basis too large for one grid");
        const int blocks = (int)((threads + 255) / 256);
        auto go = [&](auto nm) {
            constexpr int NM = decltype(nm)::value;
            switch (kind) {
            case ONE_OVERLAP:
                hipLaunchKernelGGL((one_e_kernel<ONE_OVERLAP, NM>), dim3(blocks), dim3(256), 0, st, a);
                break;
            case ONE_KINETIC:
                hipLaunchKernelGGL((one_e_kernel<ONE_KINETIC, NM>), dim3(blocks), dim3(256), 0, st, a);
                break;
            case ONE_NUCLEAR:
                hipLaunchKernelGGL((one_e_kernel<ONE_NUCLEAR, NM>), dim3(blocks), dim3(256), 0, st, a);
                break;
            case ONE_NUCLEAR_OWN:
                hipLaunchKernelGGL((one_e_kernel<ONE_NUCLEAR_OWN, NM>), dim3(blocks), dim3(256), 0, st, a);
                break;
            default: hipLaunchKernelGGL((one_e_kernel<ONE_MULTIPOLE, NM>), dim3(blocks), dim3(256), 0, st, a); break;
            }
        };
        // recurrence tables sized for the basis: s/p/d (as before), f and g
        if (b->maxL <= 2) go(std::integral_constant<int, 4>());
        else go(std::integral_constant<int, 2 * QBX_MAXL>());
        QBX_HIP(hipGetLastError());
    }
    QBX_HIP(hipMemcpyAsync(out, dout.p, NN * sizeof(double), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
}

// Concurrent per-class timing: sum the per-class durations of a recorded
// event set (each class's stretch on its auxiliary stream plus its stretch on
// the main stream, where kernels that spill are launched) into the context's
// accumulators.
void drain_timing_set(qbx_ctx* ctx, qbx_ctx::TimingSet& T)
{
    if (!T.pending) return;
    QBX_HIP(hipEventSynchronize(T.b1));
    if (ctx->tsum.size() != (size_t)QBX_N_CLASSES) ctx->tsum.assign(QBX_N_CLASSES, 0.0);
    for (int cls = 0; cls < QBX_N_CLASSES; ++cls) {
        float l = 0.f, m = 0.f;
        QBX_HIP(hipEventElapsedTime(&l, T.beg_l[cls], T.end_l[cls]));
        QBX_HIP(hipEventElapsedTime(&m, T.beg_s[cls], T.end_s[cls]));
        ctx->tsum[cls] += (double)l + (double)m;
    }
    float tot = 0.f;
    QBX_HIP(hipEventElapsedTime(&tot, T.b0, T.b1));
    ctx->tsum_total += tot;
    ctx->tcount += 1;
    T.pending = false;
}

void jk_device(qbx_ctx* ctx, qbx_basis* b, int nset, const double* dD, double* dG, double thr, int mode,
               hipStream_t st)
{
    ensure_schwarz(b);
    ensure_tasks(b, thr, mode, ctx->mixed_cut);
    ensure_hl_tasks(b, thr);
    const size_t NN = (size_t)b->nbf * b->nbf;
    const int copies = ctx->xcd_local ? QBX_MAX_XCD : 1;
    b->d_J.ensure(NN * nset * copies);
    b->d_K.ensure(NN * nset * copies);
    zero_async(b->d_J.p, NN * nset * copies, st);
    zero_async(b->d_K.p, NN * nset * copies, st);
    // timing 1: every class on the main stream, one after another, bracketed
    // by events; timing 2: the concurrent build as it runs untimed, each
    // class bracketed by events on the stream it is launched on
    std::vector<hipEvent_t> ev;
    qbx_ctx::TimingSet* T = nullptr;
    if (ctx->timing == 1) {
        ev.resize(2 * QBX_N_CLASSES + 2);
        for (auto& e : ev) QBX_HIP(hipEventCreate(&e));
        QBX_HIP(hipEventRecord(ev[2 * QBX_N_CLASSES], st));
    } else if (ctx->timing == 2) {
        T = &ctx->tsets[ctx->tnext];
        drain_timing_set(ctx, *T);
        ctx->tnext = (ctx->tnext + 1) % (int)ctx->tsets.size();
        QBX_HIP(hipEventRecord(T->b0, st));
    }
    const double* pdmax = nullptr;
    const int ns = (int)b->shells.size();
    if (ctx->density_thr > 0.0) {
        b->d_pdmax.ensure((size_t)ns * ns);
        hipLaunchKernelGGL(pair_dmax_kernel, dim3((ns * ns + 255) / 256), dim3(256), 0, st, b->d_sh.p, ns, dD,
                           2 * nset, b->nbf, b->d_pdmax.p);
        QBX_HIP(hipGetLastError());
        pdmax = b->d_pdmax.p;
    }
    // one class's J/K kernels (and, in mixed precision, its FP32 list) on ls
    auto launch_class = [&](int cls, hipStream_t ls) {
        if (ctx->timing == 1) QBX_HIP(hipEventRecord(ev[2 * cls], st));
        if (T) {
            QBX_HIP(hipEventRecord(T->beg_l[cls], ls));
            QBX_HIP(hipEventRecord(T->beg_s[cls], st));
        }
        EriArgs a = base_args(b, cls);
        a.pdmax = pdmax;
        a.scratch_stream = (void*)st;
        a.nshell = ns;
        a.dthr = ctx->density_thr;
        a.D = dD;
        a.J = b->d_J.p;
        a.K = b->d_K.p;
        a.xcd_local = ctx->xcd_local ? 1 : 0;
        a.jk_stride = NN * nset;
        a.nset = nset;
        // twin groups where their LDS rows ([set][member][A and B][N]) fit a block
        bool twin = b->use_twin[cls] && ctx->jk_rows;
        if (twin) {
            const size_t rows = (size_t)nset * b->w64[cls].twmax *
                                (qbx_ncart(qbx_class_l[cls][0]) + qbx_ncart(qbx_class_l[cls][1])) * b->nbf *
                                sizeof(double);
            const size_t boys = (size_t)QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE * sizeof(double);
            twin = rows <= (size_t)ROW_LDS_MAX && rows + boys <= (size_t)ctx->lds_max;
        }
        auto set_tasks = [&](const TaskSet& t, const TwinSet* w) {
            a.prefix = t.prefix.p;
            a.koff = t.koff.p;
            a.cpre = t.cpre.p;
            // no K-row task lists: the one-wave-per-chunk kernels run
            a.spre = ctx->jk_rows ? t.spre.p : nullptr;
            a.nsuper = t.nsuper;
            a.tchunks = t.chunks;
            a.total = t.total;
            a.nwave = t.nwave;
            a.nbra = t.nbra;
            a.gpre = w ? w->gpre.p : nullptr;
            a.gmem = w ? w->gmem.p : nullptr;
            a.ngroup = w ? w->ngroup : 0;
            a.ntwtask = w ? w->ntask : 0;
            a.twmax = w ? w->twmax : 0;
        };
        a.lds_max = ctx->lds_max;
        a.split = ctx->staged_split ? 1 : 0;
        const TaskSet& t64 = twin ? b->s64[cls] : b->t64[cls];
        set_tasks(t64, twin ? &b->w64[cls] : nullptr);
        attach_workspace(b, cls, std::max(t64.nwave, mode == QBX_PREC_MIXED ? b->t32[cls].nwave : 0), a);
        k_jk64[cls](a, ls);
        if (mode == QBX_PREC_MIXED) {
            set_tasks(b->t32[cls], nullptr);
            k_jk32[cls](a, ls);
        }
        if (ctx->timing == 1) QBX_HIP(hipEventRecord(ev[2 * cls + 1], st));
        if (T) {
            QBX_HIP(hipEventRecord(T->end_l[cls], ls));
            QBX_HIP(hipEventRecord(T->end_s[cls], st));
        }
    };
    // Without serialized timing, the class kernels run concurrently on the
    // auxiliary streams (staged classes first, then largest work first);
    // they only meet in the J/K atomics.
    const bool fork = ctx->timing != 1;
    if (fork) {
        QBX_HIP(hipEventRecord(ctx->ev_fork, st));
        for (int k = 0; k < kStreams; ++k) QBX_HIP(hipStreamWaitEvent(ctx->aux[k], ctx->ev_fork, 0));
    }
    // quartets with an f or g shell: the generic kernel (FP64 in every mode),
    // first on the last auxiliary stream
    if (b->has_high && b->hl.total > 0) {
        GenArgs g = gen_lists(b);
        g.D = dD;
        g.J = b->d_J.p;
        g.K = b->d_K.p;
        g.xcd_local = ctx->xcd_local ? 1 : 0;
        g.jk_stride = NN * nset;
        g.nset = nset;
        g.pdmax = pdmax;
        g.nshell = ns;
        g.dthr = ctx->density_thr;
        launch_eri_generic(g, b->nm, MODE_JK, fork ? ctx->aux[kStreams - 1] : st);
    }
    std::vector<int> order(QBX_N_CLASSES);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        // staged classes first: few long-running waves each, they overlap with
        // the wide low-L kernels instead of running alone at the end; then the
        // largest estimated work first, so the short classes fill the tail;
        // the classes of qbx_class_last after all of them
        if (qbx_class_last[x] != qbx_class_last[y]) return qbx_class_last[y] != 0;
        const bool sx = qbx_class_ws[x] > 0, sy = qbx_class_ws[y] > 0;
        if (sx != sy) return sx;
        return b->class_cost[x] > b->class_cost[y];
    });
    for (int oi = 0; oi < QBX_N_CLASSES; ++oi) launch_class(order[oi], fork ? ctx->aux[oi % kStreams] : st);
    if (fork) {
        for (int k = 0; k < kStreams; ++k) {
            QBX_HIP(hipEventRecord(ctx->ev_join[k], ctx->aux[k]));
            QBX_HIP(hipStreamWaitEvent(st, ctx->ev_join[k], 0));
        }
    }
    const size_t tot = NN * nset;
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       b->d_J.p, b->d_K.p, dG, b->nbf, nset, copies);
    QBX_HIP(hipGetLastError());
    QBX_HIP(hipEventRecord(b->ev_done, st));
    b->last_stream = st;
    b->built = true;
    if (T) {
        QBX_HIP(hipEventRecord(T->b1, st));
        T->pending = true;
    }
    if (ctx->timing == 1) {
        QBX_HIP(hipEventRecord(ev[2 * QBX_N_CLASSES + 1], st));
        QBX_HIP(hipEventSynchronize(ev[2 * QBX_N_CLASSES + 1]));
        ctx->last_ms.assign(QBX_N_CLASSES, 0.f);
        for (int cls = 0; cls < QBX_N_CLASSES; ++cls)
            QBX_HIP(hipEventElapsedTime(&ctx->last_ms[cls], ev[2 * cls], ev[2 * cls + 1]));
        QBX_HIP(hipEventElapsedTime(&ctx->last_total, ev[2 * QBX_N_CLASSES], ev[2 * QBX_N_CLASSES + 1]));
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
}

// Work matrices of the purification: P, Q, T (n x n each) and aux (2n + 4).
struct PurifyWork {
    double *P, *Q, *T, *aux;
};

// D = C[:, :nocc] C[:, :nocc]^T of eigen(X' F X) by trace-correcting
// purification (device X and F, n x n row-major) on stream st, using only the
// work matrices w; returns the device pointer (one of w's matrices) holding D.
// QBX_ENOCONV when the trace or the idempotency has not settled.
const double* purify_density(int n, int nocc, const double* dX, const double* dF, const PurifyWork& w, int max_iter,
                             int32_t* iterations, double* idempotency, hipStream_t st)
{
    double* dP = w.P;
    double* dQ = w.Q;
    double* dT = w.T;
    // Fock matrix in the orthonormal basis: Fo = X' F X (into dQ)
    gemm_f64(false, false, n, n, n, 1.0, dF, n, dX, n, 0.0, nullptr, 0, dT, n, st);
    gemm_f64(true, false, n, n, n, 1.0, dX, n, dT, n, 0.0, nullptr, 0, dQ, n, st);
    // spectral interval from Gershgorin discs
    double lohi[2];
    gershgorin(dQ, n, w.aux, st);
    spectral_interval(w.aux, n, w.aux + 2 * (size_t)n, st);
    QBX_HIP(hipMemcpyAsync(lohi, w.aux + 2 * (size_t)n, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    double lo = lohi[0], hi = lohi[1];
    if (!(hi > lo)) hi = lo + 1.0;
    sp2_start(dQ, n, lo, hi, dP, st);
    // trace-correcting purification (SP2): P <- P^2 or 2P - P^2, whichever
    // moves trace(P) toward nocc; quadratic convergence to the projector
    // on the nocc lowest eigenvectors once the trace is right
    double tf[2];
    double* dtf = w.aux + 2 * (size_t)n + 2;
    auto measure = [&] {
        trace_and_frobenius(dP, n, dtf, st);
        QBX_HIP(hipMemcpyAsync(tf, dtf, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    };
    measure();
    int it = 0, tail = -1;
    const double N = (double)nocc;
    for (; it < max_iter; ++it) {
        const double tr = tf[0], fr = tf[1];
        const double err = std::fabs(tr - fr);
        if (tail < 0 && err < 1e-8 * std::max(1.0, N) && std::fabs(tr - N) < 1e-6) tail = 2;
        if (tail == 0) break;
        if (tail > 0) --tail;
        if (std::fabs(fr - N) <= std::fabs(2.0 * tr - fr - N))
            gemm_f64(false, false, n, n, n, 1.0, dP, n, dP, n, 0.0, nullptr, 0, dQ, n, st);
        else
            gemm_f64(false, false, n, n, n, -1.0, dP, n, dP, n, 2.0, dP, n, dQ, n, st);
        std::swap(dP, dQ);
        measure();
    }
    if (iterations) *iterations = it;
    if (idempotency) *idempotency = std::fabs(tf[0] - tf[1]);
    if (tail != 0 || !(std::fabs(tf[0] - N) < 1e-6))
        throw QbxError(QBX_ENOCONV, "purification did not converge (no gap at the Fermi level, or max_iter too small)");
    // D = X P X'
    gemm_f64(false, false, n, n, n, 1.0, dX, n, dP, n, 0.0, nullptr, 0, dT, n, st);
    gemm_f64(false, true, n, n, n, 1.0, dT, n, dX, n, 0.0, nullptr, 0, dQ, n, st);
    return dQ;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int qbx_last_error(char* buf, size_t len)
{
    const std::string& m = g_last_error;
    if (buf && len > 0) {
        size_t n = std::min(len - 1, m.size());
        std::memcpy(buf, m.data(), n);
        buf[n] = 0;
    }
    return (int)m.size();
}

int qbx_init(int32_t device, qbx_ctx** out)
{
    return guarded([&] {
        if (!out) throw QbxError(QBX_EINVAL, "null output pointer");
        *out = nullptr;
        int ndev = 0;
        QBX_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev)
            throw QbxError(QBX_EINVAL, "device " + std::to_string(device) + " out of range (" +
                                           std::to_string(ndev) + " devices)");
        std::unique_ptr<qbx_ctx> c(new qbx_ctx());
        c->device = device;
        QBX_HIP(hipSetDevice(device));
        QBX_HIP(hipStreamCreate(&c->stream));
        for (int k = 0; k < kStreams; ++k) {
            QBX_HIP(hipStreamCreate(&c->aux[k]));
            QBX_HIP(hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
        }
        QBX_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        std::vector<double> tab, et, rows;
        build_boys_table(tab, et, rows);
        c->boys.upload(tab, c->stream);
        c->bexp.upload(et, c->stream);
        c->boysl.upload(rows, c->stream);
        QBX_HIP(hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        // address of the fixed work array on this device (device current), once per device
        {
            if (device >= kMaxDevices) throw QbxError(QBX_EUNSUP, "device index above the library's table");
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            if (!g_scratch_addr[device]) {
                void* p = nullptr;
                QBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
                if (!p) throw QbxError(QBX_EHIP, "hipGetSymbolAddress gave no address for the work array");
                g_scratch_addr[device] = (double*)p;
            }
            c->scratch = g_scratch_addr[device];
        }
        // Per-XCD J/K accumulation needs the XCC_ID register to name each
        // block's XCD: check that it reads ids below QBX_MAX_XCD and that
        // several XCDs show up (a single-XCD device keeps one copy).
        {
            const int nblk = 4096;
            DevBuf<int> ids;
            ids.ensure(nblk);
            hipLaunchKernelGGL(xcc_probe_kernel, dim3(nblk), dim3(64), 0, c->stream, ids.p);
            QBX_HIP(hipGetLastError());
            std::vector<int> h(nblk);
            QBX_HIP(hipMemcpyAsync(h.data(), ids.p, nblk * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            QBX_HIP(hipStreamSynchronize(c->stream));
            bool seen[QBX_MAX_XCD] = {};
            bool ok = true;
            for (int v : h) {
                if (v < 0 || v >= QBX_MAX_XCD) ok = false;
                else seen[v] = true;
            }
            int n = 0;
            for (bool x : seen) n += x ? 1 : 0;
            c->n_xcd = n;
            const char* env = std::getenv("QBX_XCD_LOCAL");
            const bool off = env && env[0] == '0';
            c->xcd_local = ok && n >= 2 && !off;
        }
        {
            const char* env = std::getenv("QBX_TWINS");
            c->twins = !(env && env[0] == '0');
            const char* benv = std::getenv("QBX_KET_BUCKETS");
            c->ket_buckets = !(benv && benv[0] == '0');
            const char* renv = std::getenv("QBX_JK_ROWS");
            c->jk_rows = !(renv && renv[0] == '0');
            const char* senv = std::getenv("QBX_STAGED_SPLIT");
            c->staged_split = !(senv && senv[0] == '0');
        }
        QBX_HIP(hipStreamSynchronize(c->stream));
        *out = c.release();
    });
}

void qbx_finalize(qbx_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->boys.release();
    ctx->bexp.release();
    ctx->boysl.release();
    for (int k = 0; k < kStreams; ++k) {
        (void)hipStreamSynchronize(ctx->aux[k]);
        (void)hipStreamDestroy(ctx->aux[k]);
        (void)hipEventDestroy(ctx->ev_join[k]);
    }
    (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    for (auto& T : ctx->tsets) {
        for (int c = 0; c < QBX_N_CLASSES; ++c) {
            (void)hipEventDestroy(T.beg_l[c]);
            (void)hipEventDestroy(T.end_l[c]);
            (void)hipEventDestroy(T.beg_s[c]);
            (void)hipEventDestroy(T.end_s[c]);
        }
        (void)hipEventDestroy(T.b0);
        (void)hipEventDestroy(T.b1);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int qbx_basis_create(qbx_ctx* ctx, int32_t n_prim, const double* prim_cen, const double* prim_xpn,
                     const int32_t* prim_ang, int32_t n_bf, const int32_t* bf_ptr, const int32_t* bf_prim,
                     const double* bf_w, qbx_basis** out)
{
    return guarded([&] {
        if (!ctx || !out) throw QbxError(QBX_EINVAL, "null context or output pointer");
        *out = nullptr;
        if (n_prim <= 0 || n_bf <= 0) throw QbxError(QBX_EINVAL, "empty basis");
        if (!prim_cen || !prim_xpn || !prim_ang || !bf_ptr || !bf_prim || !bf_w)
            throw QbxError(QBX_EINVAL, "null basis array");
        if (bf_ptr[0] != 0) throw QbxError(QBX_EINVAL, "bf_ptr[0] must be 0");
        std::unique_ptr<qbx_basis> b(new qbx_basis());
        b->ctx = ctx;
        b->nbf = n_bf;
        QBX_HIP(hipSetDevice(ctx->device));
        QBX_HIP(hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming));
        // ---- group basis functions into shells ----
        struct Key {
            double c[3];
            int L;
            std::vector<double> x;
        };
        std::vector<Key> keys;
        for (int f = 0; f < n_bf; ++f) {
            const int s0 = bf_ptr[f], s1 = bf_ptr[f + 1];
            if (s1 <= s0) throw QbxError(QBX_EINVAL, "basis function " + std::to_string(f) + " has no primitive");
            const int p0 = bf_prim[s0];
            if (p0 < 0 || p0 >= n_prim) throw QbxError(QBX_EINVAL, "primitive index out of range");
            const double* c0 = prim_cen + 3 * p0;
            const int32_t* a0 = prim_ang + 3 * p0;
            for (int s = s0; s < s1; ++s) {
                const int p = bf_prim[s];
                if (p < 0 || p >= n_prim) throw QbxError(QBX_EINVAL, "primitive index out of range");
                if (prim_xpn[p] <= 0.0) throw QbxError(QBX_EINVAL, "exponents must be positive");
                for (int t = 0; t < 3; ++t)
                    if (prim_cen[3 * p + t] != c0[t] || prim_ang[3 * p + t] != a0[t])
                        throw QbxError(QBX_EINVAL, "basis function " + std::to_string(f) +
                                                       " mixes centres or Cartesian exponents");
            }
            if (a0[0] < 0 || a0[1] < 0 || a0[2] < 0) throw QbxError(QBX_EINVAL, "negative angular momentum");
            const int L = a0[0] + a0[1] + a0[2];
            if (L > QBX_MAXL)
                throw QbxError(QBX_EUNSUP, "angular momentum " + std::to_string(L) + " above the build limit " +
                                               std::to_string(QBX_MAXL));
            b->maxL = std::max(b->maxL, L);
            int comp = -1;
            for (int c = 0; c < qbx_ncart(L); ++c) {
                int ijk[3];
                qbx_cart(L, c, ijk);
                if (ijk[0] == a0[0] && ijk[1] == a0[1] && ijk[2] == a0[2]) comp = c;
            }
            // try to join an existing shell with identical centre/exponents and
            // proportional contraction weights
            int found = -1;
            double scale = 1.0;
            for (int sh = (int)b->shells.size() - 1; sh >= 0 && found < 0; --sh) {
                const HostShell& H = b->shells[sh];
                const Key& K = keys[sh];
                if (H.L != L || H.bf[comp] >= 0 || (int)K.x.size() != s1 - s0) continue;
                if (K.c[0] != c0[0] || K.c[1] != c0[1] || K.c[2] != c0[2]) continue;
                bool ok = true;
                for (int s = s0; s < s1 && ok; ++s) ok = prim_xpn[bf_prim[s]] == K.x[s - s0];
                if (!ok) continue;
                int kmax = 0;
                for (int k = 1; k < H.nprim; ++k)
                    if (std::fabs(b->pcoef[H.poff + k]) > std::fabs(b->pcoef[H.poff + kmax])) kmax = k;
                const double cref = b->pcoef[H.poff + kmax];
                if (cref == 0.0) continue;
                const double sc = bf_w[s0 + kmax] / cref;
                for (int k = 0; k < H.nprim && ok; ++k) {
                    const double w = bf_w[s0 + k], pred = sc * b->pcoef[H.poff + k];
                    ok = std::fabs(w - pred) <= 1e-13 * (std::fabs(w) + std::fabs(pred)) + 1e-300;
                }
                if (ok) {
                    found = sh;
                    scale = sc;
                }
            }
            if (found < 0) {
                HostShell H;
                for (int t = 0; t < 3; ++t) H.c[t] = c0[t];
                H.L = L;
                H.nprim = s1 - s0;
                H.poff = (int)b->pexp.size();
                for (int c = 0; c < QBX_MAXC; ++c) {
                    H.bf[c] = -1;
                    H.sc[c] = 0.0;
                }
                Key K;
                for (int t = 0; t < 3; ++t) K.c[t] = c0[t];
                K.L = L;
                for (int s = s0; s < s1; ++s) {
                    b->pexp.push_back(prim_xpn[bf_prim[s]]);
                    b->pcoef.push_back(bf_w[s]);
                    K.x.push_back(prim_xpn[bf_prim[s]]);
                }
                b->shells.push_back(H);
                keys.push_back(K);
                found = (int)b->shells.size() - 1;
                scale = 1.0;
            }
            b->shells[found].bf[comp] = f;
            b->shells[found].sc[comp] = scale;
        }
        // ---- shell pairs, oriented so that L_A >= L_B ----
        const int ns = (int)b->shells.size();
        for (int s1 = 0; s1 < ns; ++s1)
            for (int s2 = 0; s2 <= s1; ++s2) {
                int A = s1, B = s2;
                if (b->shells[A].L < b->shells[B].L) std::swap(A, B);
                const HostShell& HA = b->shells[A];
                const HostShell& HB = b->shells[B];
                const int cl = pair_class_index(HA.L, HB.L);
                PairClass& P = b->pc[cl];
                const double AB[3] = {HA.c[0] - HB.c[0], HA.c[1] - HB.c[1], HA.c[2] - HB.c[2]};
                const double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
                P.shA.push_back(A);
                P.shB.push_back(B);
                P.ppo.push_back((int)b->pp.size());
                int kept = 0;
                for (int u = 0; u < HA.nprim; ++u)
                    for (int v = 0; v < HB.nprim; ++v) {
                        const double a = b->pexp[HA.poff + u], bb = b->pexp[HB.poff + v];
                        const double p = a + bb;
                        const double kp = b->pcoef[HA.poff + u] * b->pcoef[HB.poff + v] * std::exp(-a * bb / p * r2) / p;
                        if (std::fabs(kp) < QBX_PRIM_PAIR_CUT) continue;
                        ++kept;
                        PrimPair q;
                        q.p = p;
                        q.inv2p = 0.5 / p;
                        q.Px = (a * HA.c[0] + bb * HB.c[0]) / p;
                        q.Py = (a * HA.c[1] + bb * HB.c[1]) / p;
                        q.Pz = (a * HA.c[2] + bb * HB.c[2]) / p;
                        q.PAx = q.Px - HA.c[0];
                        q.PAy = q.Py - HA.c[1];
                        q.PAz = q.Pz - HA.c[2];
                        q.Kp = kp;
                        q.pad = bb;
                        b->pp.push_back(q);
                        b->ppk.push_back(u | (v << 16));
                    }
                P.ppn.push_back(kept);
                for (int t = 0; t < 3; ++t) P.AB.push_back(AB[t]);
                P.Q.push_back(0.0);
                b->p1.push_back(make_int2(cl, P.n() - 1));
            }
        // ---- upload ----
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        std::vector<DevShell> ds(ns);
        for (int s = 0; s < ns; ++s) {
            const HostShell& H = b->shells[s];
            ds[s].x = H.c[0];
            ds[s].y = H.c[1];
            ds[s].z = H.c[2];
            ds[s].L = H.L;
            ds[s].nprim = H.nprim;
            for (int c = 0; c < QBX_MAXC; ++c) {
                ds[s].bf[c] = H.bf[c];
                ds[s].sc[c] = H.sc[c];
            }
        }
        b->d_sh.upload(ds, st);
        b->d_pexp.upload(b->pexp, st);
        b->has_high = b->maxL > QBX_GEN_LMAX;
        b->nm = 4 * b->maxL;
        b->d_pp.upload(b->pp, st);
        b->d_ppk.upload(b->ppk, st);
        std::vector<int> poff(ns);
        for (int s = 0; s < ns; ++s) poff[s] = b->shells[s].poff;
        b->d_poff.upload(poff, st);
        for (int c = 0; c < QBX_N_PAIR_CLASSES; ++c) {
            PairClass& P = b->pc[c];
            P.d_shA.upload(P.shA, st);
            P.d_shB.upload(P.shB, st);
            P.d_ppo.upload(P.ppo, st);
            P.d_ppn.upload(P.ppn, st);
            P.d_AB.upload(P.AB, st);
            P.d_Q.upload(P.Q, st);
        }
        std::vector<int> p1A, p1B, p1o, p1n;
        for (const int2& e : b->p1) {
            const PairClass& P = b->pc[e.x];
            p1A.push_back(P.shA[e.y]);
            p1B.push_back(P.shB[e.y]);
            p1o.push_back(P.ppo[e.y]);
            p1n.push_back(P.ppn[e.y]);
        }
        b->d_p1A.upload(p1A, st);
        b->d_p1B.upload(p1B, st);
        b->d_p1o.upload(p1o, st);
        b->d_p1n.upload(p1n, st);
        QBX_HIP(hipStreamSynchronize(st));
        *out = b.release();
    });
}

void qbx_basis_destroy(qbx_basis* b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->built) (void)hipEventSynchronize(b->ev_done);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int qbx_basis_info(const qbx_basis* b, int32_t* n_bf, int32_t* n_shell, int32_t* n_pair, int32_t* max_l)
{
    return guarded([&] {
        if (!b) throw QbxError(QBX_EINVAL, "null basis");
        if (n_bf) *n_bf = b->nbf;
        if (n_shell) *n_shell = (int32_t)b->shells.size();
        if (n_pair) *n_pair = (int32_t)b->p1.size();
        if (max_l) *max_l = b->maxL;
    });
}

int qbx_overlap(qbx_ctx* ctx, const qbx_basis* b, double* S)
{
    return guarded([&] { run_one_e(ctx, b, ONE_OVERLAP, 0, nullptr, nullptr, nullptr, nullptr, S); });
}

int qbx_kinetic(qbx_ctx* ctx, const qbx_basis* b, double* T)
{
    return guarded([&] { run_one_e(ctx, b, ONE_KINETIC, 0, nullptr, nullptr, nullptr, nullptr, T); });
}

int qbx_nuclear(qbx_ctx* ctx, const qbx_basis* b, int32_t n_nuc, const double* Z, const double* R, double* V)
{
    return guarded([&] { run_one_e(ctx, b, ONE_NUCLEAR, n_nuc, Z, R, nullptr, nullptr, V); });
}

int qbx_nuclear_own(qbx_ctx* ctx, const qbx_basis* b, int32_t n_nuc, const double* Z, const double* R, double* V)
{
    return guarded([&] { run_one_e(ctx, b, ONE_NUCLEAR_OWN, n_nuc, Z, R, nullptr, nullptr, V); });
}

int qbx_core_hamiltonian(qbx_ctx* ctx, const qbx_basis* b, int32_t n_nuc, const double* Z, const double* R,
                         double* H)
{
    return guarded([&] {
        check_ctx_basis(ctx, b);
        if (!H) throw QbxError(QBX_EINVAL, "null output");
        const size_t NN = (size_t)b->nbf * b->nbf;
        std::vector<double> T(NN);
        run_one_e(ctx, b, ONE_KINETIC, 0, nullptr, nullptr, nullptr, nullptr, T.data());
        run_one_e(ctx, b, ONE_NUCLEAR, n_nuc, Z, R, nullptr, nullptr, H);
        for (size_t i = 0; i < NN; ++i) H[i] = T[i] + H[i];
    });
}

int qbx_multipole(qbx_ctx* ctx, const qbx_basis* b, const double* center, const int32_t* degree, double* M)
{
    return guarded([&] {
        if (!center || !degree) throw QbxError(QBX_EINVAL, "null multipole centre or degree");
        run_one_e(ctx, b, ONE_MULTIPOLE, 0, nullptr, nullptr, center, degree, M);
    });
}

int qbx_eri_dense(qbx_ctx* ctx, const qbx_basis* bc, double* eri)
{
    return guarded([&] {
        qbx_basis* b = const_cast<qbx_basis*>(bc);
        check_ctx_basis(ctx, b);
        if (!eri) throw QbxError(QBX_EINVAL, "null output");
        if (b->nbf > 256) throw QbxError(QBX_EINVAL, "dense ERI tensor limited to 256 basis functions");
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        ensure_schwarz(b);
        const size_t N = b->nbf, N4 = N * N * N * N;
        DevBuf<double> d;
        d.ensure(N4);
        zero_async(d.p, N4, st);
        // each class's task lists (host and device) stay alive until the one
        // synchronization below, so the classes' kernels queue back to back
        struct DenseTasks {
            TaskCounts tc;
            DevBuf<int64_t> pre;
            DevBuf<int> ko, cw;
        };
        std::vector<std::unique_ptr<DenseTasks>> lists;
        for (int cls = 0; cls < QBX_N_CLASSES; ++cls) {
            auto L = std::make_unique<DenseTasks>();
            TaskCounts& tc = L->tc;
            task_counts(b, cls, 0.0, QBX_PREC_FP64, 0.0, tc);
            const int nt = tc.ntb;
            if (tc.p64[nt] == 0) continue;
            L->pre.upload(tc.p64, st);
            L->ko.upload(tc.k64, st);
            L->cw.upload(tc.w64, st);
            EriArgs a = base_args(b, cls);
            a.prefix = L->pre.p;
            a.koff = L->ko.p;
            a.cpre = L->cw.p;
            a.total = tc.p64[nt];
            a.nwave = tc.w64[nt];
            a.nbra = nt;
            attach_workspace(b, cls, a.nwave, a);
            a.eri = d.p;
            k_dense[cls](a, st);
            lists.push_back(std::move(L));
        }
        if (b->has_high) {
            ensure_hl_tasks(b, 0.0);
            GenArgs g = gen_lists(b);
            g.eri = d.p;
            launch_eri_generic(g, b->nm, MODE_DENSE, st);
        }
        QBX_HIP(hipMemcpyAsync(eri, d.p, N4 * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_jk_device(qbx_ctx* ctx, qbx_basis* b, int32_t n_set, const double* dDJ, const double* dDK, double* dG,
                  double thr, int32_t mode, void* stream)
{
    return guarded([&] {
        check_ctx_basis(ctx, b);
        if (n_set <= 0) throw QbxError(QBX_EINVAL, "n_set must be positive");
        if (!dDJ || !dDK || !dG) throw QbxError(QBX_EINVAL, "null density or output pointer");
        if (mode != QBX_PREC_FP64 && mode != QBX_PREC_MIXED) throw QbxError(QBX_EINVAL, "unknown precision mode");
        if (!(thr >= 0.0)) throw QbxError(QBX_EINVAL, "Schwarz threshold must be >= 0");
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
        order_after_last_build(b, st);
        const size_t NN = (size_t)b->nbf * b->nbf;
        // the kernels read DJ and DK from one contiguous buffer
        b->d_D.ensure(2 * NN * n_set);
        QBX_HIP(hipMemcpyAsync(b->d_D.p, dDJ, NN * n_set * sizeof(double), hipMemcpyDeviceToDevice, st));
        QBX_HIP(hipMemcpyAsync(b->d_D.p + NN * n_set, dDK, NN * n_set * sizeof(double), hipMemcpyDeviceToDevice, st));
        jk_device(ctx, b, n_set, b->d_D.p, dG, thr, mode, st);
    });
}

int qbx_jk(qbx_ctx* ctx, qbx_basis* b, int32_t n_set, const double* DJ, const double* DK, double* G, double thr,
           int32_t mode)
{
    return guarded([&] {
        check_ctx_basis(ctx, b);
        if (n_set <= 0) throw QbxError(QBX_EINVAL, "n_set must be positive");
        if (!DJ || !DK || !G) throw QbxError(QBX_EINVAL, "null density or output pointer");
        if (mode != QBX_PREC_FP64 && mode != QBX_PREC_MIXED) throw QbxError(QBX_EINVAL, "unknown precision mode");
        if (!(thr >= 0.0)) throw QbxError(QBX_EINVAL, "Schwarz threshold must be >= 0");
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        order_after_last_build(b, st);
        const size_t NN = (size_t)b->nbf * b->nbf;
        b->d_D.ensure(2 * NN * n_set);
        b->d_G.ensure(NN * n_set);
        QBX_HIP(hipMemcpyAsync(b->d_D.p, DJ, NN * n_set * sizeof(double), hipMemcpyHostToDevice, st));
        QBX_HIP(hipMemcpyAsync(b->d_D.p + NN * n_set, DK, NN * n_set * sizeof(double), hipMemcpyHostToDevice, st));
        jk_device(ctx, b, n_set, b->d_D.p, b->d_G.p, thr, mode, st);
        QBX_HIP(hipMemcpyAsync(G, b->d_G.p, NN * n_set * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_density_purify(qbx_ctx* ctx, int32_t n, int32_t nocc, const double* X, const double* F, double* D,
                       int32_t max_iter, int32_t* iterations, double* idempotency)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (n <= 0 || nocc < 0 || nocc > n) throw QbxError(QBX_EINVAL, "need 0 <= nocc <= n and n > 0");
        if (!X || !F || !D) throw QbxError(QBX_EINVAL, "null matrix pointer");
        if (max_iter <= 0) throw QbxError(QBX_EINVAL, "max_iter must be positive");
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t nn = (size_t)n * n;
        ctx->la_X.ensure(nn);
        ctx->la_F.ensure(nn);
        ctx->la_P.ensure(nn);
        ctx->la_Q.ensure(nn);
        ctx->la_T.ensure(nn);
        ctx->la_aux.ensure(2 * (size_t)n + 4);
        QBX_HIP(hipMemcpyAsync(ctx->la_X.p, X, nn * sizeof(double), hipMemcpyHostToDevice, st));
        QBX_HIP(hipMemcpyAsync(ctx->la_F.p, F, nn * sizeof(double), hipMemcpyHostToDevice, st));
        const PurifyWork w{ctx->la_P.p, ctx->la_Q.p, ctx->la_T.p, ctx->la_aux.p};
        const double* dD = purify_density(n, nocc, ctx->la_X.p, ctx->la_F.p, w, max_iter, iterations, idempotency, st);
        QBX_HIP(hipMemcpyAsync(D, dD, nn * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_density_purify_noalloc(qbx_ctx* ctx, int32_t n, int32_t nocc, const double* dX, const double* dF, double* dD,
                               int32_t max_iter, int32_t* iterations, double* idempotency, void* stream)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (n <= 0 || nocc < 0 || nocc > n) throw QbxError(QBX_EINVAL, "need 0 <= nocc <= n and n > 0");
        if (n > QBX_PURIFY_NOALLOC_MAX_N)
            throw QbxError(QBX_EINVAL, "qbx_density_purify_noalloc takes n <= " +
                                           std::to_string(QBX_PURIFY_NOALLOC_MAX_N) + " (its fixed work array); got " +
                                           std::to_string(n));
        if (!dX || !dF || !dD) throw QbxError(QBX_EINVAL, "null matrix pointer");
        if (max_iter <= 0) throw QbxError(QBX_EINVAL, "max_iter must be positive");
        if (!ctx->scratch) throw QbxError(QBX_EINTERNAL, "the fixed work array has no device address");
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
        // the work matrices in the library's fixed device array (its address on
        // this device, taken at qbx_init); every entry is written before it is read
        const size_t nn = (size_t)n * n;
        double* base = ctx->scratch;
        const PurifyWork w{base, base + nn, base + 2 * nn, base + 3 * nn};
        const double* dOut = purify_density(n, nocc, dX, dF, w, max_iter, iterations, idempotency, st);
        QBX_HIP(hipMemcpyAsync(dD, dOut, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_fock_solve(qbx_ctx* ctx, int32_t n, int32_t nocc, const double* X, const double* F, double* C, double* eps,
                   double* D)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (n <= 0 || nocc < 0 || nocc > n) throw QbxError(QBX_EINVAL, "need 0 <= nocc <= n and n > 0");
        if (!X || !F || !C) throw QbxError(QBX_EINVAL, "null matrix pointer");
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        if (!ctx->blas && rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
            ctx->blas = nullptr;
            throw QbxError(QBX_EINTERNAL, "rocblas_create_handle failed");
        }
        if (rocblas_set_stream(ctx->blas, st) != rocblas_status_success)
            throw QbxError(QBX_EINTERNAL, "rocblas_set_stream failed");
        const size_t nn = (size_t)n * n;
        ctx->la_X.ensure(nn);
        ctx->la_F.ensure(nn);
        ctx->la_P.ensure(nn);
        ctx->la_Q.ensure(nn);
        ctx->la_T.ensure(nn);
        ctx->la_aux.ensure(3 * (size_t)n + 4);
        double* dX = ctx->la_X.p;
        double* dF = ctx->la_F.p;
        double* dP = ctx->la_P.p;
        double* dQ = ctx->la_Q.p;
        double* dT = ctx->la_T.p;
        double* dE = ctx->la_aux.p;             // eigenvalues (n)
        double* dOff = dE + n;                  // off-diagonal work of the tridiagonal form (n)
        double* dSgn = dOff + n;                // column signs (n)
        rocblas_int* dInfo = (rocblas_int*)(dSgn + n);
        QBX_HIP(hipMemcpyAsync(dX, X, nn * sizeof(double), hipMemcpyHostToDevice, st));
        QBX_HIP(hipMemcpyAsync(dF, F, nn * sizeof(double), hipMemcpyHostToDevice, st));
        // Fock matrix in the orthonormal basis: X' F X, symmetrized
        gemm_f64(false, false, n, n, n, 1.0, dF, n, dX, n, 0.0, nullptr, 0, dT, n, st);
        gemm_f64(true, false, n, n, n, 1.0, dX, n, dT, n, 0.0, nullptr, 0, dP, n, st);
        symmetrize_square(dP, n, st);
        // eigenpairs, ascending; the column-major eigenvector matrix V read
        // row-major is V^T
        if (rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, n, dP, n, dE, dOff, dInfo) !=
            rocblas_status_success)
            throw QbxError(QBX_EINTERNAL, "rocsolver_dsyevd failed");
        rocblas_int info = 0;
        QBX_HIP(hipMemcpyAsync(&info, dInfo, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
        if (info != 0) throw QbxError(QBX_ENOCONV, "the symmetric eigensolver did not converge");
        // C = X V with every column's first entry made non-negative
        gemm_f64(false, true, n, n, n, 1.0, dX, n, dP, n, 0.0, nullptr, 0, dQ, n, st);
        fix_column_signs(dQ, n, dSgn, st);
        QBX_HIP(hipMemcpyAsync(C, dQ, nn * sizeof(double), hipMemcpyDeviceToHost, st));
        if (eps) QBX_HIP(hipMemcpyAsync(eps, dE, n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (D) {
            // D = C[:, :nocc] C[:, :nocc]^T
            if (nocc > 0) {
                gemm_f64(false, true, n, n, nocc, 1.0, dQ, n, dQ, n, 0.0, nullptr, 0, dT, n, st);
                QBX_HIP(hipMemcpyAsync(D, dT, nn * sizeof(double), hipMemcpyDeviceToHost, st));
            } else {
                std::memset(D, 0, nn * sizeof(double));
            }
        }
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_mo_transform(qbx_ctx* ctx, int32_t rank, int32_t n_ao, int32_t n_mo, const double* ints, const double* C,
                     double* out)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (rank != 2 && rank != 4) throw QbxError(QBX_EINVAL, "rank must be 2 (one-body) or 4 (two-body)");
        if (n_ao <= 0 || n_mo <= 0) throw QbxError(QBX_EINVAL, "n_ao and n_mo must be positive");
        if (!ints || !C || !out) throw QbxError(QBX_EINVAL, "null array pointer");
        // quarter q maps [p][rest] (p of n_ao, rest of rows[q]) to [rest][i]
        size_t rows[4], big = 0;
        for (int q = 0; q < rank; ++q) {
            size_t r = 1;
            for (int t = 0; t < rank - 1; ++t) r *= (t < rank - 1 - q) ? (size_t)n_ao : (size_t)n_mo;
            if (r > (size_t)INT32_MAX) throw QbxError(QBX_EUNSUP, "tensor too large for the transform");
            rows[q] = r;
            big = std::max(big, r * std::max((size_t)n_ao, (size_t)n_mo));
        }
        size_t n_in = 1, n_out = 1;
        for (int t = 0; t < rank; ++t) {
            n_in *= (size_t)n_ao;
            n_out *= (size_t)n_mo;
        }
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf<double> buf[2], dC;
        buf[0].ensure(big);
        buf[1].ensure(big);
        dC.upload(std::vector<double>(C, C + (size_t)n_ao * n_mo), st);
        QBX_HIP(hipMemcpyAsync(buf[0].p, ints, n_in * sizeof(double), hipMemcpyHostToDevice, st));
        for (int q = 0; q < rank; ++q) {
            const int M = (int)rows[q];
            gemm_f64(true, false, M, n_mo, n_ao, 1.0, buf[q & 1].p, M, dC.p, n_mo, 0.0, nullptr, 0,
                     buf[(q + 1) & 1].p, n_mo, st);
        }
        QBX_HIP(hipMemcpyAsync(out, buf[rank & 1].p, n_out * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_mo_coulomb_ab(qbx_ctx* ctx, int32_t n_ao, int32_t n1, int32_t n2, const double* eri, const double* C1,
                      const double* C2, double* Jab)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (n_ao <= 0 || n1 <= 0 || n2 <= 0) throw QbxError(QBX_EINVAL, "n_ao, n1 and n2 must be positive");
        if (!eri || !C1 || !C2 || !Jab) throw QbxError(QBX_EINVAL, "null array pointer");
        const size_t nn = (size_t)n_ao * n_ao;
        if (nn > (size_t)INT32_MAX) throw QbxError(QBX_EUNSUP, "tensor too large for the transform");
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf<double> dE, dC1, dC2, K1, K2, P, dJ;
        dE.ensure(nn * nn);
        QBX_HIP(hipMemcpyAsync(dE.p, eri, nn * nn * sizeof(double), hipMemcpyHostToDevice, st));
        dC1.upload(std::vector<double>(C1, C1 + (size_t)n_ao * n1), st);
        dC2.upload(std::vector<double>(C2, C2 + (size_t)n_ao * n2), st);
        K1.ensure(nn * n1);
        K2.ensure(nn * n2);
        pair_columns(dC1.p, n_ao, n1, K1.p, st);
        pair_columns(dC2.p, n_ao, n2, K2.p, st);
        // P[m][cd] = sum_ab K1[ab][m] eri[ab][cd];  Jab = P K2
        P.ensure((size_t)n1 * nn);
        dJ.ensure((size_t)n1 * n2);
        const int NN = (int)nn;
        gemm_f64(true, false, n1, NN, NN, 1.0, K1.p, n1, dE.p, NN, 0.0, nullptr, 0, P.p, NN, st);
        gemm_f64(false, false, n1, n2, NN, 1.0, P.p, NN, K2.p, n2, 0.0, nullptr, 0, dJ.p, n2, st);
        QBX_HIP(hipMemcpyAsync(Jab, dJ.p, (size_t)n1 * n2 * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int32_t qbx_n_classes(void) { return QBX_N_CLASSES; }

int qbx_class_table(int32_t* l4)
{
    return guarded([&] {
        if (!l4) throw QbxError(QBX_EINVAL, "null output");
        for (int c = 0; c < QBX_N_CLASSES; ++c)
            for (int t = 0; t < 4; ++t) l4[4 * c + t] = qbx_class_l[c][t];
    });
}

int qbx_jk_stats(qbx_ctx* ctx, qbx_basis* b, double thr, double cut, int64_t* shq, int64_t* cartq, int64_t* primq,
                 int64_t* fp64q)
{
    return guarded([&] {
        check_ctx_basis(ctx, b);
        QBX_HIP(hipSetDevice(ctx->device));
        ensure_schwarz(b);
        for (int cls = 0; cls < QBX_N_CLASSES; ++cls) {
            TaskCounts tc;
            task_counts(b, cls, thr, cut > 0.0 ? QBX_PREC_MIXED : QBX_PREC_FP64, cut, tc);
            const PairClass& PX = pair_view(b, cls, 0);
            const PairClass& PY = pair_view(b, cls, 1);
            const int nx = PX.n();
            // prefix sums over kets (sorted order) of primitive-pair counts and of
            // unique basis-function pairs (n(n+1)/2 for a diagonal shell pair)
            const int* l = qbx_class_l[cls];
            auto upairs = [&](const PairClass& P, int k, int la, int lb) -> int64_t {
                const int64_t na = qbx_ncart(la), nb = qbx_ncart(lb);
                return P.shA[k] == P.shB[k] ? na * (na + 1) / 2 : na * nb;
            };
            std::vector<int64_t> kp(PY.n() + 1, 0), kf(PY.n() + 1, 0);
            for (int j = 0; j < PY.n(); ++j) {
                kp[j + 1] = kp[j] + PY.ppn[j];
                kf[j + 1] = kf[j] + upairs(PY, j, l[2], l[3]);
            }
            int64_t sq = 0, pq = 0, s64 = 0, cq = 0;
            for (int i = 0; i < nx; ++i)
                for (int k = 0; k < PY.nbk; ++k) {
                    const int t = i * PY.nbk + k;
                    const int64_t n64 = tc.p64[t + 1] - tc.p64[t];
                    const int64_t n32 = tc.p32[t + 1] - tc.p32[t];
                    const int64_t cnt = n64 + n32;
                    const int s0 = PY.bstart[k];
                    sq += cnt;
                    s64 += n64;
                    pq += (int64_t)PX.ppn[i] * (kp[s0 + cnt] - kp[s0]);
                    const int64_t fi = upairs(PX, i, l[0], l[1]);
                    const int64_t fk = kf[s0 + cnt] - kf[s0];
                    // the diagonal quartet (bra == ket, the last ket of its range) counts fi(fi+1)/2
                    if (X_is_Y(cls) && i >= s0 && i < s0 + cnt)
                        cq += fi * (fk - fi) + fi * (fi + 1) / 2;
                    else
                        cq += fi * fk;
                }
            if (shq) shq[cls] = sq;
            if (cartq) cartq[cls] = cq;
            if (primq) primq[cls] = pq;
            if (fp64q) fp64q[cls] = s64;
        }
    });
}

int qbx_set_timing(qbx_ctx* ctx, int32_t mode)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (mode < 0 || mode > 2) throw QbxError(QBX_EINVAL, "timing mode must be 0, 1 or 2");
        QBX_HIP(hipSetDevice(ctx->device));
        if (mode == 2 && ctx->tsets.empty()) {
            ctx->tsets.resize(32);
            for (auto& T : ctx->tsets) {
                for (int c = 0; c < QBX_N_CLASSES; ++c) {
                    QBX_HIP(hipEventCreate(&T.beg_l[c]));
                    QBX_HIP(hipEventCreate(&T.end_l[c]));
                    QBX_HIP(hipEventCreate(&T.beg_s[c]));
                    QBX_HIP(hipEventCreate(&T.end_s[c]));
                }
                QBX_HIP(hipEventCreate(&T.b0));
                QBX_HIP(hipEventCreate(&T.b1));
                T.pending = false;
            }
        }
        if (mode == 2) {
            // start a fresh average
            for (auto& T : ctx->tsets) T.pending = false;
            ctx->tsum.assign(QBX_N_CLASSES, 0.0);
            ctx->tsum_total = 0.0;
            ctx->tcount = 0;
        }
        ctx->timing = mode;
    });
}

int qbx_last_timing(qbx_ctx* ctx, float* ms, float* total)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        QBX_HIP(hipSetDevice(ctx->device));
        if (ctx->timing == 2) {
            // drain in recording order: the oldest pending set sits at tnext
            const int n = (int)ctx->tsets.size();
            for (int k = 0; k < n; ++k) drain_timing_set(ctx, ctx->tsets[(ctx->tnext + k) % n]);
            if (ctx->tcount == 0) throw QbxError(QBX_EINVAL, "no timed qbx_jk call yet");
            if (ms)
                for (int c = 0; c < QBX_N_CLASSES; ++c) ms[c] = (float)(ctx->tsum[c] / ctx->tcount);
            if (total) *total = (float)(ctx->tsum_total / ctx->tcount);
            ctx->tsum.assign(QBX_N_CLASSES, 0.0);
            ctx->tsum_total = 0.0;
            ctx->tcount = 0;
            return;
        }
        if (ctx->last_ms.empty()) throw QbxError(QBX_EINVAL, "no timed qbx_jk call yet");
        if (ms)
            for (int c = 0; c < QBX_N_CLASSES; ++c) ms[c] = ctx->last_ms[c];
        if (total) *total = ctx->last_total;
    });
}

int qbx_boys_device(qbx_ctx* ctx, int32_t order, int32_t n, const double* T, double* F)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (order < 0 || order > 8) throw QbxError(QBX_EINVAL, "Boys order must be in [0, 8]");
        if (n <= 0) throw QbxError(QBX_EINVAL, "n must be positive");
        if (!T || !F) throw QbxError(QBX_EINVAL, "null argument or output");
        for (int i = 0; i < n; ++i)
            if (!(T[i] >= 0.0 && T[i] <= 1e6)) throw QbxError(QBX_EINVAL, "Boys arguments must lie in [0, 1e6]");
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf<double> dT, dF;
        dT.upload(std::vector<double>(T, T + n), st);
        dF.ensure((size_t)n * (order + 1));
        boys_probe(ctx->boysl.p + (size_t)order * QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE, order, dT.p, n, dF.p, st);
        QBX_HIP(hipMemcpyAsync(F, dF.p, (size_t)n * (order + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_prim_one_e(qbx_ctx* ctx, int32_t kind, int32_t n, const double* A, const double* a, const int32_t* la,
                   const double* B, const double* b, const int32_t* lb, const double* C, const int32_t* degree,
                   double* out)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (kind < QBX_PRIM_OVERLAP || kind > QBX_PRIM_MULTIPOLE) throw QbxError(QBX_EINVAL, "unknown integral kind");
        if (n < 0) throw QbxError(QBX_EINVAL, "n must be non-negative");
        if (n == 0) return;
        if (!A || !a || !la || !B || !b || !lb || !out) throw QbxError(QBX_EINVAL, "null argument or output");
        if ((kind == QBX_PRIM_COULOMB || kind == QBX_PRIM_MULTIPOLE) && !C) throw QbxError(QBX_EINVAL, "null point");
        PrimOneArgs g{};
        g.n = n;
        g.kind = kind == QBX_PRIM_OVERLAP ? ONE_OVERLAP
                 : kind == QBX_PRIM_KINETIC ? ONE_KINETIC
                 : kind == QBX_PRIM_COULOMB ? ONE_NUCLEAR
                                            : ONE_MULTIPOLE;
        if (kind == QBX_PRIM_MULTIPOLE) {
            if (!degree) throw QbxError(QBX_EINVAL, "null multipole degree");
            for (int t = 0; t < 3; ++t) {
                if (degree[t] < 0 || degree[t] > 4) throw QbxError(QBX_EUNSUP, "multipole degree per axis must be in [0, 4]");
                g.deg[t] = degree[t];
            }
        }
        if (C)
            for (int t = 0; t < 3; ++t) g.C[t] = C[t];
        for (int i = 0; i < n; ++i) {
            if (!(a[i] > 0.0) || !(b[i] > 0.0)) throw QbxError(QBX_EINVAL, "exponents must be positive");
            int sa = 0, sb = 0;
            for (int t = 0; t < 3; ++t) {
                if (la[3 * i + t] < 0 || lb[3 * i + t] < 0) throw QbxError(QBX_EINVAL, "negative angular momentum");
                sa += la[3 * i + t];
                sb += lb[3 * i + t];
            }
            if (sa > kPrimLMax || sb > kPrimLMax)
                throw QbxError(QBX_EUNSUP, "primitive angular momentum above " + std::to_string(kPrimLMax));
        }
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf<double> dA, da, dB, db, dout;
        DevBuf<int> dla, dlb;
        dA.upload(std::vector<double>(A, A + 3 * (size_t)n), st);
        da.upload(std::vector<double>(a, a + n), st);
        dB.upload(std::vector<double>(B, B + 3 * (size_t)n), st);
        db.upload(std::vector<double>(b, b + n), st);
        dla.upload(std::vector<int>(la, la + 3 * (size_t)n), st);
        dlb.upload(std::vector<int>(lb, lb + 3 * (size_t)n), st);
        dout.ensure(n);
        g.A = dA.p;
        g.a = da.p;
        g.B = dB.p;
        g.b = db.p;
        g.la = dla.p;
        g.lb = dlb.p;
        g.boys = ctx->boys.p;
        g.bexp = ctx->bexp.p;
        g.out = dout.p;
        hipLaunchKernelGGL(prim_one_e_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g);
        QBX_HIP(hipGetLastError());
        QBX_HIP(hipMemcpyAsync(out, dout.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_prim_eri(qbx_ctx* ctx, int32_t n, const double* cen, const double* xpn, const int32_t* ang, double* out)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (n < 0) throw QbxError(QBX_EINVAL, "n must be non-negative");
        if (n == 0) return;
        if (!cen || !xpn || !ang || !out) throw QbxError(QBX_EINVAL, "null argument or output");
        for (int i = 0; i < n; ++i) {
            int lt = 0;
            for (int k = 0; k < 4; ++k) {
                if (!(xpn[4 * i + k] > 0.0)) throw QbxError(QBX_EINVAL, "exponents must be positive");
                int lk = 0;
                for (int t = 0; t < 3; ++t) {
                    if (ang[12 * i + 3 * k + t] < 0) throw QbxError(QBX_EINVAL, "negative angular momentum");
                    lk += ang[12 * i + 3 * k + t];
                }
                if (lk > kPrimLMax)
                    throw QbxError(QBX_EUNSUP, "primitive angular momentum above " + std::to_string(kPrimLMax));
                lt += lk;
            }
            if (lt > QBX_PRIM_ERI_LTOT)
                throw QbxError(QBX_EUNSUP, "total angular momentum of a primitive quartet above " +
                                               std::to_string(QBX_PRIM_ERI_LTOT));
        }
        std::lock_guard<std::mutex> lk(ctx->mu);
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DevBuf<double> dc, dx, dout;
        DevBuf<int> da;
        dc.upload(std::vector<double>(cen, cen + 12 * (size_t)n), st);
        dx.upload(std::vector<double>(xpn, xpn + 4 * (size_t)n), st);
        da.upload(std::vector<int>(ang, ang + 12 * (size_t)n), st);
        dout.ensure(n);
        prim_eri(n, dc.p, dx.p, da.p, ctx->boys.p, ctx->bexp.p, dout.p, st);
        QBX_HIP(hipMemcpyAsync(out, dout.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    });
}

int qbx_set_density_screening(qbx_ctx* ctx, double thr)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (!(thr >= 0.0)) throw QbxError(QBX_EINVAL, "density screening threshold must be >= 0");
        ctx->density_thr = thr;
    });
}

int qbx_basis_shells(const qbx_basis* b, int32_t* n_shell, int32_t* n_radial, int32_t* shell_L,
                     int32_t* shell_nprim, int32_t* shell_poff, int32_t* shell_bf, double* shell_sc,
                     double* shell_cen, double* radial_xpn, double* radial_coef)
{
    return guarded([&] {
        if (!b) throw QbxError(QBX_EINVAL, "null basis");
        const int ns = (int)b->shells.size();
        if (n_shell) *n_shell = ns;
        if (n_radial) *n_radial = (int32_t)b->pexp.size();
        for (int s = 0; s < ns; ++s) {
            const HostShell& H = b->shells[s];
            if (shell_L) shell_L[s] = H.L;
            if (shell_nprim) shell_nprim[s] = H.nprim;
            if (shell_poff) shell_poff[s] = H.poff;
            for (int c = 0; c < QBX_MAXC; ++c) {
                if (shell_bf) shell_bf[QBX_MAXC * s + c] = H.bf[c];
                if (shell_sc) shell_sc[QBX_MAXC * s + c] = H.sc[c];
            }
            for (int t = 0; t < 3; ++t)
                if (shell_cen) shell_cen[3 * s + t] = H.c[t];
        }
        for (size_t k = 0; k < b->pexp.size(); ++k) {
            if (radial_xpn) radial_xpn[k] = b->pexp[k];
            if (radial_coef) radial_coef[k] = b->pcoef[k];
        }
    });
}

}  // extern "C"

namespace {

// RHF (Da = Db = null) or UHF energy gradient; see qbx_rhf_gradient and
// qbx_uhf_gradient in qbx.h.
void hf_gradient(qbx_ctx* ctx, qbx_basis* b, const double* Dt, const double* Da, const double* Db, const double* W,
                 int32_t n_nuc, const double* Z, const double* R, double thr, double* g_xpn, double* g_lin,
                 double* g_cen)
{
    {
        check_ctx_basis(ctx, b);
        if (!Dt || !W || !g_xpn || !g_lin || !g_cen) throw QbxError(QBX_EINVAL, "null density or output pointer");
        if ((Da == nullptr) != (Db == nullptr)) throw QbxError(QBX_EINVAL, "UHF needs both spin densities");
        const bool uhf = Da != nullptr;
        if (n_nuc <= 0 || !Z || !R) throw QbxError(QBX_EINVAL, "the gradient needs at least one nucleus");
        if (!(thr >= 0.0)) throw QbxError(QBX_EINVAL, "Schwarz threshold must be >= 0");
        if (b->maxL > 1) throw QbxError(QBX_EUNSUP, "the gradient kernels cover s and p shells");
        for (const HostShell& H : b->shells)
            if (H.nprim > QBX_GRAD_MAXP)
                throw QbxError(QBX_EUNSUP, "the gradient kernels take at most " + std::to_string(QBX_GRAD_MAXP) +
                                               " primitives per shell");
        QBX_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        ensure_schwarz(b);
        ensure_tasks(b, thr, QBX_PREC_FP64, 0.0);
        const size_t NN = (size_t)b->nbf * b->nbf;
        const int ns = (int)b->shells.size(), nr = (int)b->pexp.size();
        DevBuf<double> dD, dW, dZ, dR, gx, gl, gc;
        {
            // [D] for RHF, [D, D_alpha, D_beta] for UHF
            std::vector<double> h(Dt, Dt + NN);
            if (uhf) {
                h.insert(h.end(), Da, Da + NN);
                h.insert(h.end(), Db, Db + NN);
            }
            dD.upload(h, st);
        }
        dW.upload(std::vector<double>(W, W + NN), st);
        dZ.upload(std::vector<double>(Z, Z + n_nuc), st);
        dR.upload(std::vector<double>(R, R + 3 * n_nuc), st);
        gx.ensure(nr);
        gl.ensure(nr);
        gc.ensure(3 * (size_t)ns);
        zero_async(gx.p, nr, st);
        zero_async(gl.p, nr, st);
        zero_async(gc.p, 3 * (size_t)ns, st);
        // one-electron part
        OneGradArgs o;
        std::memset(&o, 0, sizeof(o));
        o.sh = b->d_sh.p;
        o.pp = b->d_pp.p;
        o.ppk = b->d_ppk.p;
        o.pA = b->d_p1A.p;
        o.pB = b->d_p1B.p;
        o.po = b->d_p1o.p;
        o.pn = b->d_p1n.p;
        o.npair = (int)b->p1.size();
        o.N = b->nbf;
        o.boys = ctx->boys.p;
        o.bexp = ctx->bexp.p;
        o.nnuc = n_nuc;
        o.Z = dZ.p;
        o.R = dR.p;
        o.Dt = dD.p;
        o.W = dW.p;
        o.poff = b->d_poff.p;
        o.g_xpn = gx.p;
        o.g_lin = gl.p;
        o.g_cen = gc.p;
        const int64_t threads = (int64_t)o.npair * kGradPairPrims;
        if (threads > 0) {
            hipLaunchKernelGGL(one_e_grad_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, o);
            QBX_HIP(hipGetLastError());
        }
        // two-electron part
        for (int cls = 0; cls < QBX_N_CLASSES; ++cls) {
            const TaskSet& t64 = b->t64[cls];
            if (t64.nwave <= 0) continue;
            EriLaunch f = grad_launcher(cls);
            if (!f) throw QbxError(QBX_EINTERNAL, "no gradient kernel for an occupied class");
            EriArgs a = base_args(b, cls);
            const int lt = qbx_class_l[cls][0] + qbx_class_l[cls][1] + qbx_class_l[cls][2] + qbx_class_l[cls][3];
            a.boysl = ctx->boysl.p + (size_t)(lt + 1) * QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE;
            a.D = dD.p;
            a.nset = uhf ? 2 : 1;
            a.prefix = t64.prefix.p;
            a.koff = t64.koff.p;
            a.cpre = t64.cpre.p;
            a.total = t64.total;
            a.nwave = t64.nwave;
            a.nbra = t64.nbra;
            a.ppk = b->d_ppk.p;
            a.shell_poff = b->d_poff.p;
            a.g_xpn = gx.p;
            a.g_lin = gl.p;
            a.g_cen = gc.p;
            f(a, st);
        }
        QBX_HIP(hipMemcpyAsync(g_xpn, gx.p, nr * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipMemcpyAsync(g_lin, gl.p, nr * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipMemcpyAsync(g_cen, gc.p, 3 * (size_t)ns * sizeof(double), hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
    }
}

}  // namespace

extern "C" {

int qbx_rhf_gradient(qbx_ctx* ctx, qbx_basis* b, const double* Dt, const double* W, int32_t n_nuc,
                     const double* Z, const double* R, double thr, double* g_xpn, double* g_lin, double* g_cen)
{
    return guarded([&] { hf_gradient(ctx, b, Dt, nullptr, nullptr, W, n_nuc, Z, R, thr, g_xpn, g_lin, g_cen); });
}

int qbx_uhf_gradient(qbx_ctx* ctx, qbx_basis* b, const double* Dt, const double* Da, const double* Db,
                     const double* W, int32_t n_nuc, const double* Z, const double* R, double thr, double* g_xpn,
                     double* g_lin, double* g_cen)
{
    return guarded([&] {
        if (!Da || !Db) throw QbxError(QBX_EINVAL, "null spin density");
        hf_gradient(ctx, b, Dt, Da, Db, W, n_nuc, Z, R, thr, g_xpn, g_lin, g_cen);
    });
}

const char* qbx_source_digest(void)
{
#ifdef QBX_SOURCE_DIGEST
    return QBX_SOURCE_DIGEST;
#else
    return "unknown";
#endif
}

int qbx_set_mixed_cut(qbx_ctx* ctx, double cut)
{
    return guarded([&] {
        if (!ctx) throw QbxError(QBX_EINVAL, "null context");
        if (!(cut >= 0.0)) throw QbxError(QBX_EINVAL, "cut must be >= 0");
        ctx->mixed_cut = cut;
    });
}

}  // extern "C"
