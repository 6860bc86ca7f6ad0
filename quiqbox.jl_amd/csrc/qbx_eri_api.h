// qbx_eri_api.h -- argument block and launcher declaration of the ERI kernels.
#pragma once
#include "qbx_device.h"
#include <hip/hip_runtime.h>
#include <cstdint>

namespace qbx {
enum EriMode { MODE_JK = 0, MODE_DENSE = 1, MODE_SCHWARZ = 2 };

// J/K kernel with the bra's K rows in LDS: ROW_WAVES waves per block, up to
// ROW_CHUNKS chunks of 64 kets of one bra per block, at most ROW_LDS_MAX bytes
// of rows (larger bases use the one-wave-per-chunk kernel).
#ifndef QBX_ROW_WAVES
#define QBX_ROW_WAVES 4
#endif
constexpr int ROW_WAVES = QBX_ROW_WAVES;
// Chunks per task of the low-L K-row and twin kernels: fewer, larger tasks
// (fewer K-row flushes) measured 50.9 -> 49.6 ms per Fock build of 24 H2O in
// cc-pVDZ at 32 instead of 16, the 592-function cluster unchanged; at 64 the
// cluster lost (38.6 -> 39.4 ms).
constexpr int ROW_CHUNKS = 32;
// Chunks per task of the staged K-row kernel: its waves are long, and smaller
// tasks balance better (592-function cluster: (dp|pp) 5.05 -> 4.33 ms, build
// 41.0 -> 40.1 ms against 16; water's low-L classes lost 4% at 8).
constexpr int STAGED_ROW_CHUNKS = 8;
constexpr int ROW_LDS_MAX = 64 * 1024;
// Twin bras: shell pairs of one pair class with the same primitive geometry
// and primitive pairs, differing only in contraction coefficients (shells
// that share centre, L and exponents, as the 1s and 2s contractions of
// cc-pVDZ oxygen).  A group of up to TWIN_MAX of them is evaluated together
// by classes with at most TWIN_MAX_ACC accumulators.
constexpr int TWIN_MAX = 4;
constexpr int TWIN_WAVES = 8;   // waves per block of the twin kernel (its K rows take up to 61 KB of LDS)
constexpr int TWIN_MAX_ACC = 9;

struct PairClassDev {
    int n;
    const int* shA;
    const int* shB;
    const int* ppo;
    const int* ppn;
    const double* AB;
    const double* Q;    // Schwarz bound per pair (sorted descending)
};

struct EriArgs {
    PairClassDev bra, ket;
    const PrimPair* pp;
    const DevShell* sh;
    const double* boys;
    const double* bexp;
    const double* boysl;      // per-order Boys rows of the class's total L
    const int64_t* prefix;    // [nbra+1] prefix over bras of the screened ket counts
    const int* koff;          // [nbra] first ket of each bra
    const int* cpre;          // [nbra+1] prefix over bras of ceil(kets / 64): waves
    const int* spre;          // [nbra+1] prefix over bras of ceil(waves / tchunks): K-row blocks
    int nsuper;               // number of K-row blocks
    int tchunks;              // chunks per K-row task of the staged kernels (spre)
    const int* gpre;          // [ngroup+1] twin groups: prefix of tasks (ROW_CHUNKS chunks of the union range)
    const int* gmem;          // [ngroup][TWIN_MAX][3]: bra, first ket, end of kets (bra < 0: no member)
    int ngroup;               // twin groups of this class (0: none)
    int ntwtask;              // their tasks
    int twmax;                // largest group (LDS rows per set)
    int nbk;                  // ket buckets per bra: task t is bra t / nbk, bucket t % nbk
    int lds_max;              // LDS bytes a block may use on this device
    int64_t total;            // Schwarz mode: number of diagonal pairs
    int nwave;                // J/K and dense modes: number of waves
    int nbra;
    int same;
    const double* D;   // [2*nset][N][N]: DJ sets then DK sets
    double* J;         // [copies][nset][N][N]
    double* K;
    // xcd_local: each XCD accumulates into its own copy of J and K (copy =
    // the XCC_ID of the wave, stride jk_stride doubles) with workgroup-scope
    // atomics, which the XCD's L2 resolves; otherwise one copy, device scope
    int xcd_local;
    size_t jk_stride;
    int nset;
    int N;
    double* eri;       // dense mode, N^4
    double* qout;      // Schwarz mode, per bra pair
    // density screening (J/K mode): skip quartets with Q_ab Q_cd max|D| < dthr,
    // max|D| over the six shell-pair blocks that the quartet reads
    const double* pdmax;  // [n_shell][n_shell], NULL: off
    int nshell;
    double dthr;
    void* scratch_stream;  // main stream: kernels with scratch are launched there
    // staged classes: workspace of ws_slots wave slots, each
    // EriClass::NE * NCD * 64 values of the kernel's arithmetic type
    void* ws;
    int ws_slots;
    int split;                // classes with Cls::SPLIT > 1: share each wave task among SPLIT waves (0: off)
    // gradient kernels: radial indices (k_a | k_b << 16) of every primitive pair,
    // first radial primitive of every shell, and the outputs: per radial
    // primitive dE/d(exponent) and c dE/dc, per shell dE/d(centre)
    const int* ppk;
    const int* shell_poff;
    double* g_xpn;
    double* g_lin;
    double* g_cen;
};


// Defined (explicitly instantiated) in the generated per-class translation units.
template <int LA, int LB, int LC, int LD, typename R, int MODE>
void launch_eri(const EriArgs& a, hipStream_t s);

// Gradient kernels (classes whose shells are all s or p), generated units eri_<class>_grad.hip.
template <int LA, int LB, int LC, int LD>
void launch_eri_grad(const EriArgs& a, hipStream_t s);

typedef void (*EriLaunch)(const EriArgs&, hipStream_t);

// Shell quartets with a shell above the generated classes (f and g): the
// generic kernel of qbx_eri_generic.hip.  Bras are the pairs of the higher
// pair classes in descending Schwarz order; the kets of bra i are the first
// nlow[i] of the lower pairs (descending Schwarz order) followed by the first
// prefix[i+1] - prefix[i] - nlow[i] higher pairs (those up to bra i itself).
// One wave per quartet, lanes over its Cartesian component quartets.
struct GenArgs {
    const DevShell* sh;
    const PrimPair* pp;
    const int* ppk;           // radial indices (k_a | k_b << 16) of every primitive pair
    const int* poff;          // first radial primitive of every shell
    const double* pexp;       // radial exponents
    const double* boys;       // one-electron Boys table (F_0..F_32 on the 0.1 grid)
    const double* bexp;
    int nbra;
    int64_t total;            // quartets (waves)
    const int64_t* prefix;    // [nbra + 1]
    const int* nlow;          // [nbra]
    const int *hA, *hB, *ho, *hn;  // higher pairs: shells, primitive-pair offset and count
    const double* hQ;
    const int *lA, *lB, *lo, *ln;  // lower pairs
    const double* lQ;
    int N;
    // J/K mode (as EriArgs)
    const double* D;
    double* J;
    double* K;
    int xcd_local;
    size_t jk_stride;
    int nset;
    const double* pdmax;
    int nshell;
    double dthr;
    // dense mode
    double* eri;
    // Schwarz mode: every bra with itself, max |(ab|ab)| as the bits of a
    // non-negative double
    unsigned long long* qbits;
};

// nm: the per-axis recurrence size (4 x the highest shell angular momentum
// of the quartets, 12 or 16); mode: MODE_JK, MODE_DENSE or MODE_SCHWARZ
void launch_eri_generic(const GenArgs& a, int nm, int mode, hipStream_t s);

// Primitive-quartet ERIs with unit coefficients (qbx_eri_generic.hip), one
// thread each: cen [n][4][3], xpn [n][4], ang [n][4][3] (order a, b, c, d of
// (ab|cd)), total angular momentum at most QBX_PRIM_ERI_LTOT.
#define QBX_PRIM_ERI_LTOT 24
void prim_eri(int n, const double* cen, const double* xpn, const int* ang, const double* boys, const double* bexp,
              double* out, hipStream_t s);

// The ERI kernels' Boys function alone (qbx_boys_probe.hip): F_0..F_L at n
// device arguments T from the order-L table rows, into F (n x (L+1)).
void boys_probe(const double* rows, int L, const double* T, int n, double* F, hipStream_t s);
}  // namespace qbx
