// engine_internal.hpp -- the engine context (device tables + walker state)
// shared by the kernel translation units and the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/champ_gpu.h"
#include "device_common.hpp"

namespace champ {

struct DmcState;  // host side of a DMC population (engine_dmc.hip)

struct EcpParams {
    int lpot[CHAMP_MAXCT];
    int nterm[CHAMP_MAXCT][CHAMP_MPS_L];
    int power[CHAMP_MAXCT][CHAMP_MPS_L][CHAMP_MGAUSS];
    double coef[CHAMP_MAXCT][CHAMP_MPS_L][CHAMP_MGAUSS];
    double expo[CHAMP_MAXCT][CHAMP_MPS_L][CHAMP_MGAUSS];
    double znuc[CHAMP_MAXCT];
    double wq[CHAMP_MPS_QUAD];
    int nquad, nloc;
};

// Per-kernel-class timing with HIP events on the context stream.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
};

enum KClass { K_MOVE_A = 0, K_MOVE_GEMM, K_MOVE_B, K_HPSI_JAS, K_HPSI_AO, K_HPSI_GEMM, K_HPSI_SLATER,
              K_HPSI_ECP, K_HPSI_NONLOC, K_HPSI_PSI, K_HPSI_ELOC, K_MD, K_OTHER, K_NCLASS };

}  // namespace champ

struct champ_ctx {
    // ---- dimensions
    int W = 0, NE = 0, NUP = 0, NDN = 0, NM = 0, NC = 0, NCT = 0, NB = 0, NO = 0, NQ = 0;
    int nloc = 0, prec = 0, hpsi_chunk = 0, pair_cap = 0;
    double pecent = 0, cdet0 = 1;
    double tau = 0.5;  // time step of the metropolis kernels, set by each metrop1 / metrop1_batch call
    int imetro = 1;    // mover of the sweep kernels: 1 metrop1 (drift-diffusion), 6 metrop6 (accelerated)
    double m6_deltar = 5.0, m6_deltat = 1.0, m6_fbias = 1.0;  // metrop6 parameters (parser.f90:402-404)
    hipStream_t stream = nullptr;
    int dev = 0;                 // device the context was created on
    double *scratch = nullptr;   // that device's address of the static scratch of champ_gpu_basis_fns_noalloc

    // ---- tables
    double *cent = nullptr;          // [NC][3]
    int *iwctype = nullptr;          // [NC]
    champ::SplineTab tab{};
    double *rmax_ao = nullptr;       // [NB] rmaxwf of each AO's shell
    // radial instances: one per (centre, radial function of its type)
    int NRI = 0;
    int *ri_cent = nullptr, *ri_type = nullptr, *ri_rad = nullptr;  // [NRI]
    double *ri_rmax = nullptr;                                      // [NRI]
    int *ao_ri = nullptr;                                           // [NB] radial instance of each AO
    int *ri_perm = nullptr;   // [NRI] instances in decreasing order of their AO count
    int *ri_aoff = nullptr;   // [NRI+1] offsets of each instance's AOs in ri_aol
    int *ri_aol = nullptr;    // [NB] AOs grouped by radial instance
    double *gref = nullptr;                                         // [nr][2] grid node radii
    int lu_every = 0;  // moves between FP64 recomputes of the Slater inverse inside a sweep (0: only in hpsi)
    int *ao_cent = nullptr, *ao_type = nullptr, *ao_rad = nullptr, *ao_slm = nullptr;  // [NB]
    double *coef = nullptr;          // [NO][NB]
    float *coef32 = nullptr;         // [NO][NB] FP32 copy for the mixed path's per-move contraction
    double *cocc[2] = {nullptr, nullptr};  // [nel][NB] coef rows of occupied orbitals
    int *iworbd = nullptr;           // [NE]
    champ::JasParams hjas{};
    champ::JasParams *jas = nullptr;
    champ::EcpParams hecp{};
    champ::EcpParams *ecp = nullptr;
    double *xq0 = nullptr, *yq0 = nullptr, *zq0 = nullptr;  // [NQ]
    std::vector<double> h_xq0, h_yq0, h_zq0;

    // ---- walker state
    double *xT = nullptr;      // [W][3][NE]
    double *r_en = nullptr;    // [W][NC][NE]
    double *rr_en = nullptr;   // [W][NC][NE]  scale_dist (A/B scaling)
    double *fso = nullptr;     // [W][NE][NE]  symmetric, diagonal = e-n
    double *g1 = nullptr;      // [W][NE][3][NE] g1[i][k][j] = grad_i of pair(i,j)
    double *g2 = nullptr;      // [W][NE][3][NE] g2[i][k][j] = grad_j of pair(i,j)
    double *vjT = nullptr;     // [W][3][NE]
    double *fsumo = nullptr;   // [W]
    double *d2o = nullptr;     // [W]
    double *dorb = nullptr;    // [W][NE][3][NO]
    double *slmi = nullptr;    // [W][2][NM][NM]
    double *lu_ws = nullptr;   // [2 hpsi_chunk][NM (NM+1)]: Slater work matrices too large for LDS
    double *detiab = nullptr;  // [W][2]
    double *ddx = nullptr;     // [W][NE][3]
    double *d2dx2 = nullptr;   // [W][NE]
    double *psido = nullptr, *psijo = nullptr, *psi2o = nullptr, *eold = nullptr, *ekin = nullptr;
    double *pe_local = nullptr, *vpsp = nullptr /*[W][2]*/;
    int *status = nullptr;
    uint64_t *rng = nullptr;   // [W][4]
    double *xq = nullptr, *yq = nullptr, *zq = nullptr;  // [W][NQ]

    // ---- move scratch
    double *xnew = nullptr;    // [W][3]
    double *dfus2o = nullptr;  // [W]
    double *rnew = nullptr, *rrnew = nullptr;  // [W][NC]
    double *fsn_row = nullptr; // [W][NE]
    double *g1n = nullptr, *g2n = nullptr;     // [W][3][NE]
    double *vjnT = nullptr;    // [W][3][NE]
    double *fsumn = nullptr;   // [W]
    double *aomove = nullptr;  // [W][4][NB] (the mixed path stores FP32 values in its first half)
    double *orbn = nullptr;    // [W][4][NO]
    double *ratio = nullptr, *psidn = nullptr, *psijn = nullptr, *vnew = nullptr /*[W][3]*/, *vold = nullptr;
    double *m6rec = nullptr;   // [W][M6_REC] metrop6 proposal record

    // ---- DMC (dmc_ps): per-walker state of the population (allocated by
    // champ_gpu_dmc_init; W is the capacity, nwalk the live walkers, which
    // occupy slots 0 .. nwalk-1) and the host-side population control
    double *dmc_acc = nullptr;   // [W][DMC_ACC] move accumulators of the current step
    int *dmc_age = nullptr;      // [W] steps since the walker last moved (iage)
    double *dmc_wt = nullptr;    // [W] weights
    double *dmc_fratio = nullptr;  // [W] velocity ratio of the last step (fratio)
    double *dmc_eprev = nullptr; // [W] local energy of the previous step
    double *dmc_out = nullptr;   // [W][DMC_OUT] per-walker terms of the step's estimators
    int dmc_on = 0, dmc_W = 0;   // DMC initialised; capacity (the context's walker count at init)
    champ::DmcState *dmc = nullptr;

    // ---- parameter derivatives (engine_deriv.hip): d ln psi / d p of the
    // current walkers, parameter-major [nparm][W], and the SR sums' scratch
    double *der_ot = nullptr, *der_s = nullptr, *der_v = nullptr;
    int der_cap = 0, der_scap = 0;
    // parameter map (champ_gpu_set_deriv_map): Mt [nout][nraw] over the raw
    // parameters of der_map_kinds, and the mapped O [nout][W]
    double *der_map = nullptr, *der_om = nullptr;
    // the map's unit columns are copies: der_src [nout] holds the raw
    // parameter a column copies, or -(1 + d) for the d-th dense column, whose
    // coefficients der_map holds ([der_nd][nraw]); der_dt [der_nd][W]
    int *der_src = nullptr;
    double *der_dt = nullptr;
    int der_nd = 0, der_dt_cap = 0;
    // orbital terms (champ_gpu_set_orb_terms): [nterm][2] (io, jo) and the slot
    // of each orbital in the reference determinant per spin [2][NO] (-1: empty)
    int *der_orb = nullptr, *der_occ = nullptr;
    int der_norb = 0;
    int der_map_kinds = 0, der_map_nraw = 0, der_map_nout = 0, der_om_cap = 0;
    // SR sums kept on the device (champ_gpu_sr_accumulate): S [n][n], [O | E O]
    // [2n], the conjugate-gradient vectors [2n]; weight and E_L totals
    double *sr_acc = nullptr, *sr_accv = nullptr, *sr_z = nullptr;
    // scratch of the split-K GEMM: two M x N partials (gemm_reserve_splitk)
    double *gemm_tmp = nullptr;
    size_t gemm_tmp_cap = 0;
    int sr_n = 0, sr_kinds = 0;
    double sr_wsum = 0.0, sr_esum = 0.0;
    unsigned char *acc = nullptr;
    double *esum_dev = nullptr;  // [4] accepted, esum, esum2, nflag

    // ---- hpsi scratch
    double *aoall = nullptr;   // [chunk][NE][5][NB]
    double *orball = nullptr;  // [chunk][NE][5][NO]
    double *wmat = nullptr;    // [W][NE][NB]
    int *pair_count = nullptr, *pair_off = nullptr, *pair_total = nullptr;  // [W], [W+1], [2]
    int *grp_off = nullptr;                    // [W+1] prefix sums of the per-walker pair groups
    int *pair_e = nullptr, *pair_c = nullptr;  // [W*pair_cap]
    double *pair_vps = nullptr;                // [W*pair_cap][MPS_L]
    double *pair_val = nullptr;                // [W*pair_cap]
    double *pair_psi = nullptr;                // [W*pair_cap][NQ] psi ratio at each quadrature point
    float2 *enp = nullptr;                     // mixed path: [W][NC][NE] partner pair table (enp_pack)
    int tstore = 0;                            // hpsi keeps each quadrature point's t_vpsp in pair_psi (T-moves)
    double *rrT = nullptr;                     // [W][NE][NC] rr_en with the centre innermost
    int *nxquad = nullptr;                     // [W]

    // ---- multideterminant (ndet > 1): excitation tables per spin and state
    int md = 0, md_ndet = 0, md_NV = 0, md_NU = 0, md_nu[2] = {0, 0}, md_nv[2] = {0, 0};
    int *md_du = nullptr, *md_dd = nullptr;   // [ndet] unique spin determinant of each det
    double *md_c = nullptr;                   // [ndet] cdet with the reordering phase
    double *md_ph = nullptr;                  // [ndet] the reordering phase (+-1) of each determinant
    std::vector<double> md_hph;               // host copy of md_ph (coefficient updates)
    std::vector<int> md_hwdet[2];             // determinant of each w-matrix entry per spin (md_wcf order)
    int *md_rank[2] = {nullptr, nullptr}, *md_hole[2] = {nullptr, nullptr}, *md_part[2] = {nullptr, nullptr};  // [nu][4]
    int *md_virt[2] = {nullptr, nullptr};     // [nv] orbital of each particle index
    int *md_woff[2] = {nullptr, nullptr}, *md_woth[2] = {nullptr, nullptr};  // dets of each spin determinant (CSR):
    double *md_wcf[2] = {nullptr, nullptr};                                  // other spin's determinant, c_k
    int *md_yoff[2] = {nullptr, nullptr}, *md_yent[2] = {nullptr, nullptr};  // cofactors gathered per Y cell
    int *md_kslot[2] = {nullptr, nullptr}, *md_vidx[2] = {nullptr, nullptr};  // [NO] slot / particle index or -1
    int *md_zoff[2] = {nullptr, nullptr}, *md_zent[2] = {nullptr, nullptr};  // (u, b) of each particle (CSR)
    double *md_aa = nullptr;    // [W][2][2 buffers][NM][NV]
    double *md_D = nullptr;     // [W][2][2][NU]
    double *md_tb = nullptr;    // [W][NU][4] scratch: C_u^T times a column of the inverse
    double *md_G = nullptr;     // [W][NU][16] hpsi scratch: weighted cofactors
    double *md_S = nullptr, *md_Sn = nullptr;  // [W] psi_D / (D_up,ref D_dn,ref), current and proposed
    int *md_cur = nullptr;      // [W][2] current buffer per spin
    double *md_vcol = nullptr, *md_vcoln = nullptr;  // [W][NO] column iel of V, current and proposed
    double *md_V = nullptr;     // [W][NE][NO]
    double *md_Y = nullptr, *md_Z = nullptr;  // [W][NV][NM] scratch
    double *md_Wt = nullptr;    // [W][NU] scratch
    double *md_g = nullptr;     // [W][NV] scratch

    // ---- choices of the autotuner (champ_gpu_autotune): GEMM kind per call site
    // (move orbitals, hpsi orbitals, ECP coefficient rows) and quadrature form
    int tune_gemm[3] = {-1, -1, -1};
    int tune_nonloc = -1;  // -1: shape rule, 0: matrix cores, 1: VALU form

    // ---- metrop1 sweep as a HIP graph (captured per tau); off by default:
    // replaying the ~500-node sweep graph measured 5% slower than launching
    int graphs = 0;
    hipGraphExec_t sweep_exec = nullptr;
    double sweep_key[5] = {0, 0, 0, 0, 0};  // tau, mover and metrop6 parameters the graph was captured with

    // ---- timing
    bool timing = false;
    std::vector<champ::Timer> ev_pool;
    std::vector<std::pair<int, int>> ev_used;  // (class, pool index)
    double class_ms[champ::K_NCLASS] = {0};
    long long class_n[champ::K_NCLASS] = {0};
};

namespace champ {
// DMC move accumulators per walker (engine_kernels.hip k_move_b, engine_dmc.hip)
// DMC_FNORMO: the all-electron mover's forward normalisation of the current move
enum { DMC_PP = 0, DMC_DFUS2AC, DMC_DFUS2UN, DMC_PSUM, DMC_NODECR, DMC_IACCEPT, DMC_FNORMO, DMC_ACC = 8 };
// per-walker estimator terms of a DMC step
enum { DMC_O_WT = 0, DMC_O_WE, DMC_O_PE, DMC_O_TPB, DMC_O_TAU, DMC_O_PSUM, DMC_O_NODECR, DMC_OUT = 8 };
// kernel launchers (engine_kernels.hip)
int launch_move_a(champ_ctx *c, int iel, int mode);
int launch_move_gemm(champ_ctx *c);
int launch_move_b(champ_ctx *c, int iel, int mode);
int launch_drift_old(champ_ctx *c, int iel);
int launch_stay(champ_ctx *c, int iel);
int launch_hpsi(champ_ctx *c);
int launch_slater_refresh(champ_ctx *c);
int launch_rotqua(champ_ctx *c);
int launch_basis_points(champ_ctx *c, int npt, const double *dpos, double *dout);
// kind: GEMM_AUTO (shape rule), GEMM_TILE (16 x 16 per wavefront), GEMM_STRIP
// (16 x 80 per wavefront, B chunks in LDS), GEMM_VALU, GEMM_F32 (the mixed
// path's FP32 move contraction) or GEMM_SPLITK (the tile form over the two
// halves of K into a reserved scratch, then their sum in a fixed order)
enum { GEMM_AUTO = -1, GEMM_TILE = 0, GEMM_STRIP = 1, GEMM_VALU = 2, GEMM_F32 = 3, GEMM_SPLITK = 4 };
int gemm_reserve_splitk(champ_ctx *c, int M, int N);
int launch_gemm(champ_ctx *c, bool bt, int M, int N, int K, const double *A, int lda, long long sA, const double *B,
                int ldb, long long sB, double *Cm, int ldc, long long sC, int batch, int kind = GEMM_AUTO);
int launch_wmat(champ_ctx *c);
// FP32 A [M][K] times FP32 B^T (B stored [N][K]) into FP64 C (matrix cores)
int launch_syrk_lower(champ_ctx *c, int n, int K, const double *A, int lda, double *Cm, int ldc);
int launch_gemm_f32(champ_ctx *c, int M, int N, int K, const float *A, int lda, const float *B, int ldb, double *Cm,
                    int ldc);
void dmc_free(champ_ctx *c);
void deriv_free(champ_ctx *c);
int md_init(champ_ctx *c, const champ_system_t *s);
void md_free(champ_ctx *c);
int md_set_cdet(champ_ctx *c, const double *cdet);
int launch_orbital_values(champ_ctx *c, int w0, int nw);
int launch_md_hpsi(champ_ctx *c, int w0, int nw);
int launch_md_pre(champ_ctx *c, int iel);
int launch_md_post(champ_ctx *c, int iel);
void set_error(const std::string &s);
void tbegin(champ_ctx *c, int cls, int *slot);
void tend(champ_ctx *c, int slot);
}  // namespace champ

#define HIPCHECK(x)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            champ::set_error(std::string(#x) + ": " + hipGetErrorString(e_));             \
            return CHAMP_ERR_HIP;                                                         \
        }                                                                                 \
    } while (0)
