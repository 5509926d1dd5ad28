// sw2dg_adjoint.hip -- 4D-Var layer on the MI355X: the reference's
// differentiable ("_E") forward and its hand-written adjoint.
//
// Reference path (assim/swSpE_checkpoint.cpp):
//   build_checkpoint_records (97-455)  -> assim_record: dry run that fixes dt,
//       the per-series BC values and the observation triggers of every step,
//       landing on observation times, and keeps a state checkpoint every
//       `stride` steps (sqrt(N) checkpointing)
//   ckpt_segment_E (584-715)           -> k_e_update (fluxes + update) replayed with
//       the frozen schedule; cost at observation steps
//   get_spatial_gradient_checkpointed (722-1281) -> assim_gradient: reverse
//       loop over segments (K-1 .. window start), each segment recomputed
//       forward from its checkpoint, then reversed step by step by
//       hand-written adjoint kernels (Enzyme's role in the reference); time
//       tiling resets the adjoint at tile tops (:932-970)
//   replay_trajectory_cost (1311-1433) -> assim_cost
// Controls: porosity per cell (the reference's control), friction
// coefficients per cell (Strickler K; Manning zones map to them on the host)
// and boundary values per step and series (hydrograph controls).
// All gathers run in fixed reference order: no atomics, deterministic.
#include <hip/hip_runtime.h>

#include <limits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"
#include "sp_adjoint.h"
#include "sp_physics.h"
#include "sw2dg_internal.h"

using namespace sw2dg_detail;
using sp::kG;
using sp::kSmall;

namespace sw2dg_detail {

struct EClock {
    int oi;                  // next observation index (dry run)
    int n_obs;
    int n_series;
    double t_max, dt_max, cfl;
};

// the dry run's gauge sampling, done by the update (gcell = nullptr: none):
// the depth at each gauge cell after a step that reaches an observation
// time (trig[n] >= 0), the series a twin experiment observes
struct GaugeSample {
    const int *trig;
    const int *gcell;
    int n_g;
    double *gser;
};

struct Controls {
    const double *phi, *K1, *K2;
    // each internal edge side's factor edge_coeff(length, area, phi) of its
    // cell, at 2 e + side (so also at a cell's incidence entry); computed
    // once per control upload (k_coe) and read by the Godunov update and the
    // edge reverse (the MUSCL-EVR edge pass and stored update, with VALU to
    // spare, form it in place: the two-sided array gathers worse than elen)
    const double *coe;
};

struct AssimState {
    // observations (device order)
    int n_obs = 0, n_gauges = 0, normalize = 0;
    bool frames = false;
    std::vector<double> obs_t;
    double *d_obs_t = nullptr;
    int *d_gcell = nullptr;
    double *d_gw = nullptr, *d_gobs = nullptr;
    double *d_gser = nullptr;  // [n_obs][n_gauges] depths sampled by the last dry run
    double *d_oframes = nullptr;  // [n_obs][3 nc] states at the observation steps (opt-in)
    bool keep_frames = false;
    int *d_gfirst = nullptr, *d_gnext = nullptr;  // per-cell gauge chains (gauge order)
    double *d_fh = nullptr, *d_fqx = nullptr, *d_fqy = nullptr, *d_wh = nullptr, *d_wq = nullptr;
    // schedule
    int stride = 0, n_tiles = 1;
    double t_window = -1e300;
    int64_t max_steps = 0, n_steps = 0;
    double *d_T = nullptr, *d_dt = nullptr, *d_bc = nullptr;
    std::vector<double> h_dt;  // host copy of the recorded dt sequence (replay launches read it)
    int *d_trig = nullptr;
    EClock *d_ck = nullptr;
    bool recorded = false;
    // states: checkpoints and one segment of trajectory, layout [h | qx | qy] per state
    double *d_ckpt = nullptr;  // (n_ckpt) x 3nc
    int64_t n_ckpt = 0;
    double *d_seg = nullptr;   // (stride + 1) x 3nc
    // work
    double *d_Fseg = nullptr;  // stride x 3 n_bnd: boundary fluxes of every replayed step
    double *d_F = nullptr;     // MUSCL-EVR dry run: 5 n_int + 3 n_bnd edge fluxes of the step
    double *d_sw = nullptr;    // MUSCL-EVR dry run: 2 n_int swept terms
    double *d_Pseg = nullptr;  // stride x 6 nc: balanced state + internal increments per step
    double *d_sxy = nullptr;   // 2 nc: limited free-surface slopes of the step (MUSCL-EVR)
    double *d_Sseg = nullptr;  // stride x 2 nc: the slopes of every step of the replayed segment
    // resident trajectory (Godunov): the dry run inside a gradient keeps, for
    // the steps of each segment it has a block for, what the replay of that
    // segment would recompute bit for bit (states, balanced states + internal
    // increments, boundary fluxes), so the reverse sweep reads it in place of
    // replaying.  One block per segment, allocated on first use while the
    // blocks take at most 60 % of the device memory free at the first one.
    int resident_mode = 1;            // 0 never, 1 when the blocks fit
    std::vector<double *> res_blk;    // per segment: [(stride + 1) x 3nc | stride x 6nc | stride x 3nb]
    std::vector<char> res_kept;       // per segment: the last gradient's dry run filled its block
    size_t res_budget = 0, res_bytes = 0;
    int64_t seg_blk = -1;             // >= 0: the seg_* views address this segment's block
    double *d_partial = nullptr;
    double *d_J = nullptr;     // [1 + n_obs]: total, then the cost of each observation
    double *d_cpart = nullptr; // frame-cost block partials
    // adjoint
    double *d_lam = nullptr, *d_lams = nullptr, *d_part = nullptr, *d_side = nullptr;
    double *d_sd = nullptr;    // MUSCL-EVR: 4 n_int slope adjoints per edge side
    double *d_zn = nullptr;    // MUSCL-EVR: 2 n_int neighbour free-surface adjoints (per edge side)
    double *d_gphi = nullptr, *d_gK1 = nullptr, *d_gK2 = nullptr, *d_gbce = nullptr,
           *d_gbc = nullptr;
    // change of variable of the porosity control (phi_correct / dphi_dm,
    // swSpE.h:59-107): 0 none, 1 clamp [0, 1.05], 2 logistic (lambda), 3
    // normalise (m - lo) / (hi - lo) with lo / hi over the free cells, frozen
    // at each dry run (update_phi_norm_bounds, swSpE_core.cpp:229-266)
    int phi_mode = 0;
    double phi_lambda = 1.0, norm_lo = 0.0, norm_scale = 1.0;
    std::vector<int64_t> free_idx;
    bool mask_wall = false;
    std::vector<double> m_last;  // the control passed with the last dry run / cost (reference order)
    // explosion guard (swSpE_checkpoint.cpp:1035-1117): ratio (<= 0: off),
    // gradient snapshot at the last milestone, truncation of the last gradient
    double expl_ratio = 1e8;
    int64_t trunc_seg = -1;
    int64_t forced_seg = -1;    // frozen truncation window (forced_first_seg_), -1 = dynamic
    double *d_wm = nullptr;     // 3 nc: gphi | gK1 | gK2 at the watermark
    double *d_nrm = nullptr;    // block partials of ||lam_h||^2, then the sum
    double dh_peak = 0.0;       // largest ||lam_h|| at a segment start in the last gradient
    // controls (device order)
    double *d_phi = nullptr, *d_K1 = nullptr, *d_K2 = nullptr;
    double *d_coe = nullptr;  // [2 n_int] edge-side factors of the current phi
    // series -> boundary edges (renumbered) in reference order
    int *d_ser_ptr = nullptr, *d_ser_edges = nullptr;
    // optional per-launch timing (sw2dg_assim_profile): an event pair per launch
    hipStream_t stream = nullptr;
    bool prof = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    std::vector<std::pair<const char *, size_t>> marks;
    std::vector<void *> allocs;
    template <class T>
    T *alloc(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        if (hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        return (T *)p;
    }
    ~AssimState() {
        for (void *p : allocs)
            if (p) (void)hipFree(p);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

void assim_destroy(AssimState *a) { delete a; }

// event pair around one launch when profiling; -1 otherwise
inline long prof_begin(AssimState *a) {
    if (!a->prof) return -1;
    while (a->ev.size() < a->ev_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        a->ev.push_back(e);
    }
    const size_t i = a->ev_used;
    a->ev_used += 2;
    (void)hipEventRecord(a->ev[i], a->stream);
    return (long)i;
}
inline void prof_end(AssimState *a, long i, const char *name) {
    if (i < 0) return;
    (void)hipEventRecord(a->ev[i + 1], a->stream);
    a->marks.emplace_back(name, (size_t)i);
}
#define KLAUNCH(A, KERN, ...)                       \
    do {                                            \
        const long pk_ = prof_begin(A);             \
        hipLaunchKernelGGL(KERN, __VA_ARGS__);      \
        prof_end(A, pk_, #KERN);                    \
    } while (0)
// the same with the name given (kernels whose template arguments need parentheses)
#define KLAUNCH_NAMED(A, NAME, KERN, ...)           \
    do {                                            \
        const long pk_ = prof_begin(A);             \
        hipLaunchKernelGGL(KERN, __VA_ARGS__);      \
        prof_end(A, pk_, NAME);                     \
    } while (0)

}  // namespace sw2dg_detail

namespace {

__device__ __forceinline__ double smax_(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// _E forward kernels
// ---------------------------------------------------------------------------

// interpolateTimeSeries_E (assim/swSpE.h:138-153).  The bracket is the
// first it with !(t > tt[it + 1]); on a non-decreasing table that predicate
// is monotone and bisection finds the same it as the reference's scan
// (sorted = 0: the scan itself)
__device__ double interp_e(const double *tt, const double *vv, int n, double t, int sorted) {
    if (n == 0) return 0.0;
    if (n == 1 || t <= tt[0] - 1.0e-12) return vv[0];
    if (t >= tt[n - 1]) return vv[n - 1];
    int it = 0;
    if (sorted) {
        int hi = n - 1;
        while (it < hi) {
            const int mid = (it + hi) >> 1;
            if (t > tt[mid + 1]) it = mid + 1;
            else hi = mid;
        }
    } else {
        while (it < n - 1 && t > tt[it + 1]) it++;
    }
    const double t1 = tt[it], t2 = tt[it + 1], v1 = vv[it], v2 = vv[it + 1];
    const double den = t2 - t1;
    if (den < 1e-12) return v1;
    return v1 + (v2 - v1) * (t - t1) / den;
}

// the internal-edge incidences of cell i in its reference order, for the
// slope functions of sp_adjoint.h (free surface z = h + zb of the state U)
struct NbDev {
    const MeshDev &m;
    const double *U;
    int b;
    __device__ int inc(int k) const { return m.int_inc[b + k]; }
    __device__ int nbr(int k) const { return m.int_nbr[b + k]; }
    __device__ double zj(int k) const {
        const int j = nbr(k);
        return U[j] + m.zb[j];
    }
    __device__ double xj(int k) const { return m.cx[nbr(k)]; }
    __device__ double yj(int k) const { return m.cy[nbr(k)]; }
    __device__ double xm(int k) const { return m.exm[inc(k) >> 1]; }
    __device__ double ym(int k) const { return m.eym[inc(k) >> 1]; }
    __device__ int side(int k) const { return inc(k) & 1; }
};

// the same, from registers (cells of at most 4 internal edges)
struct NbReg {
    double z[4], x[4], y[4], xm_[4], ym_[4];
    int s[4];
    __device__ double zj(int k) const { return z[k]; }
    __device__ double xj(int k) const { return x[k]; }
    __device__ double yj(int k) const { return y[k]; }
    __device__ double xm(int k) const { return xm_[k]; }
    __device__ double ym(int k) const { return ym_[k]; }
    __device__ int side(int k) const { return s[k]; }
};

// MUSCL-EVR: per cell, free-surface gradient over the cell's neighbours and
// the reference's in-place limiter replayed over the cell's internal edges
// in ascending order (computeMusclSlopes, swSpHypMusclEvr.cpp:11-115; a
// cell's slope is only changed by its own edges, so the per-cell replay is
// exact).  sxy = [dz/dx | dz/dy].
__global__ __launch_bounds__(kBlock) void k_e_slopes(MeshDev m, const double *__restrict__ U,
                                                     double *__restrict__ sxy, const double *T,
                                                     int64_t n, double t_max) {
    if (T && !(T[n] < t_max)) return;
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    const int nc = m.n_cell;
    if (i >= nc) return;
    const int b = m.int_ptr[i], deg = m.int_ptr[i + 1] - b;
    const double zi = U[i] + m.zb[i], xi = m.cx[i], yi = m.cy[i];
    double sx, sy;
    if (deg >= 2 && deg <= 4) {  // triangles and quads: every incidence's data loaded once, together
        NbReg r;
        int in[4], nb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            in[j] = (j < deg) ? m.int_inc[b + j] : 0;
            nb[j] = (j < deg) ? m.int_nbr[b + j] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j < deg) {
                const int c2 = nb[j], e = in[j] >> 1;
                r.z[j] = U[c2] + m.zb[c2];
                r.x[j] = m.cx[c2];
                r.y[j] = m.cy[c2];
                r.xm_[j] = m.exm[e];
                r.ym_[j] = m.eym[e];
                r.s[j] = in[j] & 1;
            }
        }
        if (deg == 2) spe::slope_fwd(r, 2, zi, xi, yi, sx, sy);
        else if (deg == 3) spe::slope_fwd(r, 3, zi, xi, yi, sx, sy);
        else spe::slope_fwd(r, 4, zi, xi, yi, sx, sy);
    } else {
        const NbDev nb{m, U, b};
        spe::slope_fwd(nb, deg, zi, xi, yi, sx, sy);
    }
    sxy[i] = sx;
    sxy[nc + i] = sy;
}

// _E Riemann states of the two sides of internal edge e: Godunov
// (hypGodunovComputeLRStates_E, swSpE_hyp.cpp:507-573) or, MUSCL, the
// MUSCL-EVR reconstruction from the step's slopes sxy with dt_recon
// (swSpE_muscl.cpp:66-180)
template <bool MUSCL>
__device__ __forceinline__ void e_int_states(const MeshDev &m, const Opts &o, const double *__restrict__ U,
                                             const double *__restrict__ sxy, double dt_recon, int e,
                                             spe::Side &sL, spe::Side &sR) {
    const int nc = m.n_cell;
    const int L = m.eL[e], R = m.eR[e];
    const double nx = m.enx[e], ny = m.eny[e];
    if (MUSCL) {
        const double xm = m.exm[e], ym = m.eym[e];
        sL = spe::recon_muscl(0, U[L], U[nc + L], U[2 * nc + L], m.zb[L], m.beta[L], sxy[L], sxy[nc + L],
                              nx, ny, xm - m.cx[L], ym - m.cy[L], dt_recon);
        sR = spe::recon_muscl(1, U[R], U[nc + R], U[2 * nc + R], m.zb[R], m.beta[R], sxy[R], sxy[nc + R],
                              nx, ny, xm - m.cx[R], ym - m.cy[R], dt_recon);
    } else {
        sL = spe::lr_fwd(U[L], U[nc + L], U[2 * nc + L], m.zb[L], m.beta[L], nx, ny, o.h_min);
        sR = spe::lr_fwd(U[R], U[nc + R], U[2 * nc + R], m.zb[R], m.beta[R], nx, ny, o.h_min);
    }
}

// boundary state of boundary edge j of cell i
template <bool MUSCL>
__device__ __forceinline__ spe::Side e_bnd_state(const MeshDev &m, const Opts &o, const double *__restrict__ U,
                                                 int i, int j) {
    const int nc = m.n_cell;
    return MUSCL ? spe::lr_bnd_muscl(U[i], U[nc + i], U[2 * nc + i], m.zb[i], m.beta[i], m.bnx[j], m.bny[j])
                 : spe::lr_fwd(U[i], U[nc + i], U[2 * nc + i], m.zb[i], m.beta[i], m.bnx[j], m.bny[j],
                               o.h_min);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o2 = __shfl_xor(v, off, 64);
        v = (o2 < v) ? o2 : v;
    }
    return v;
}

// MUSCL-EVR: the reconstruction is costly enough that the forward computes
// each edge once and keeps its fluxes (the Godunov passes compute fluxes in
// the update itself).  Threads [0, n_int): internal edges -> F[5e..5e+4] =
// {F1, then per side the momentum contribution without dt: (mx, my) of side
// 0, (mx, my) of side 1} (side_contrib's sums, the factor of each side's
// cell applied here), and, in the dry run, the swept terms sw[2e + side];
// threads [n_int, n_int + n_bnd): boundary edges -> F[5 n_int + 3j ..] (and
// Fb, the reverse sweep's copy, when given).  Dry run (T given): the
// boundary value interpolated at T[n] and the reconstruction with the
// previous step's dt (*dtp, 0 before the first), as
// build_checkpoint_records does; replay (T null): the recorded boundary
// values bcv and the step's own dt, dt_fix.
__global__ __launch_bounds__(kBlock) void k_e_edges_muscl(MeshDev m, Opts o, Controls cs, BcDev bc,
                                                          const double *__restrict__ U,
                                                          double *__restrict__ F, double *__restrict__ sw,
                                                          const double *T, int64_t n, double t_max,
                                                          const double *__restrict__ sxy,
                                                          const double *dtp, double dt_fix,
                                                          const double *__restrict__ bcv,
                                                          double *__restrict__ Fb_out) {
    if (T && !(T[n] < t_max)) return;
    const int e = xcd_block() * kBlock + (int)threadIdx.x;
    if (e >= m.n_int) {
        const int j = e - m.n_int;
        if (j >= m.n_bnd) return;
        const int i = m.bcell[j];
        const spe::Side sd = e_bnd_state<true>(m, o, U, i, j);
        const int ser = m.bser[j];
        const double v = T ? interp_e(bc.t, bc.v + (size_t)ser * bc.n_times, bc.n_times, T[n], bc.sorted)
                           : bcv[ser];
        double f[3];
        spe::bnd_fwd(o.h_min, o.phi_min, m.beta[i], cs.phi[i], m.zb[i], sd.z, sd.q, sd.r, sd.lp,
                     bc.type[ser], v, f);
        double *Fb = F + 5 * (size_t)m.n_int + 3 * (size_t)j;
        Fb[0] = f[0];
        Fb[1] = f[1];
        Fb[2] = f[2];
        if (Fb_out) {
            Fb_out[3 * (size_t)j] = f[0];
            Fb_out[3 * (size_t)j + 1] = f[1];
            Fb_out[3 * (size_t)j + 2] = f[2];
        }
        return;
    }
    const int L = m.eL[e], R = m.eR[e];
    spe::Side sL, sR;
    e_int_states<true>(m, o, U, sxy, T ? (dtp ? *dtp : 0.0) : dt_fix, e, sL, sR);
    double f[5];
    spe::hllc_fwd(o.h_min, o.phi_min, m.beta[L], m.beta[R], cs.phi[L], cs.phi[R], m.zb[L], m.zb[R],
                  sL.z, sL.q, sL.r, sR.z, sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, f);
    const double len = m.elen[e], nx = m.enx[e], ny = m.eny[e];
    const double coL = spe::edge_coeff(len, m.area[L], cs.phi[L], o.phi_min);
    const double coR = spe::edge_coeff(len, m.area[R], cs.phi[R], o.phi_min);
    double *Fe = F + 5 * (size_t)e;
    Fe[0] = f[0];
    Fe[1] = (-coL * nx) * f[1] + (coL * nx) * f[3] + (coL * ny) * f[2];
    Fe[2] = (-coL * ny) * f[1] + (coL * ny) * f[3] + (-coL * nx) * f[2];
    Fe[3] = (coR * nx) * f[1] + (coR * nx) * f[4] + (-coR * ny) * f[2];
    Fe[4] = (coR * ny) * f[1] + (coR * ny) * f[4] + (coR * nx) * f[2];
    if (sw) {
        double m1 = fabs(sL.lm), v = fabs(sL.lp);
        if (m1 < v) m1 = v;
        if (m1 < (v = fabs(sR.lm))) m1 = v;
        if (m1 < (v = fabs(sR.lp))) m1 = v;
        sw[2 * e] = (U[L] > o.h_min) ? m1 * len : 0.0;
        sw[2 * e + 1] = (U[R] > o.h_min) ? m1 * len : 0.0;
    }
}

// the MUSCL-EVR dry run's CFL from the swept terms of k_e_edges_muscl (the
// same sums as k_e_cfl<true>)
__global__ __launch_bounds__(kBlock) void k_e_cfl_sw(MeshDev m, Opts o, const double *__restrict__ U,
                                                     const double *__restrict__ sw, const EClock *ck,
                                                     double *partial, const double *T, int64_t n) {
    if (!(T[n] < ck->t_max)) return;
    __shared__ double red[kBlock / 64];
    const int lb = xcd_block();
    const int i = lb * kBlock + (int)threadIdx.x;
    double cand = ck->dt_max;
    if (i < m.n_cell) {
        const double h = U[i];
        double swept = 0.0;
        if (h > o.h_min) {
            for (int k = m.bnd_ptr[i]; k < m.bnd_ptr[i + 1]; ++k) {
                const int j = m.bnd_inc[k];
                const spe::Side s = e_bnd_state<true>(m, o, U, i, j);
                const double lm = (fabs(s.lp) < fabs(s.lm)) ? fabs(s.lm) : fabs(s.lp);
                swept += lm * m.blen[j];
            }
        }
        const int ie = m.int_ptr[i + 1];
        for (int kb = m.int_ptr[i]; kb < ie; kb += 4) {  // four incidences' loads in flight, summed in order
            int in[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) in[j] = (kb + j < ie) ? m.int_inc[kb + j] : 0;
            double w[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) w[j] = (kb + j < ie) ? sw[in[j]] : 0.0;  // 2 e + side
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (kb + j < ie) swept += w[j];
        }
        const double den = smax_(swept, kSmall);
        const double x = ck->cfl * (m.area[i] / den);
        if (x < cand) cand = x;
    }
    cand = wave_min(cand);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = cand;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < kBlock / 64; ++w) v = (red[w] < v) ? red[w] : v;
        partial[lb] = v;
    }
}

// dry-run CFL (hypComputeTimeStep_E, swSpE_hyp.cpp:578-635) and the step
// schedule of build_checkpoint_records (swSpE_checkpoint.cpp:224-366): per
// cell, the swept term of each boundary edge (wet cell) and of each internal
// edge (max |lambda| of its two sides times its length, when this side is
// wet), summed in the cell's reference incidence order; per-block minima
template <bool MUSCL>
__global__ __launch_bounds__(kBlock) void k_e_cfl(MeshDev m, Opts o, const double *__restrict__ U,
                                                  const double *__restrict__ sxy, const double *dtp,
                                                  const EClock *ck, double *partial, const double *T,
                                                  int64_t n) {
    if (!(T[n] < ck->t_max)) return;
    __shared__ double red[kBlock / 64];
    const int lb = xcd_block();
    const int i = lb * kBlock + (int)threadIdx.x;
    const double dt_max = ck->dt_max;
    double cand = dt_max;
    const int nc = m.n_cell;
    if (i < nc) {
        const double h = U[i];
        double swept = 0.0;
        if (h > o.h_min) {
            for (int k = m.bnd_ptr[i]; k < m.bnd_ptr[i + 1]; ++k) {
                const int j = m.bnd_inc[k];
                const spe::Side s = e_bnd_state<MUSCL>(m, o, U, i, j);
                const double lm = (fabs(s.lp) < fabs(s.lm)) ? fabs(s.lm) : fabs(s.lp);
                swept += lm * m.blen[j];
            }
            // the dry run reconstructs with the previous step's dt (0 before the first)
            const double dt_recon = (MUSCL && dtp) ? *dtp : 0.0;
            if (!MUSCL) {  // own data once, neighbours through int_nbr (as e_balance)
                const double qx0 = U[nc + i], qy0 = U[2 * nc + i], zb0 = m.zb[i], be0 = m.beta[i];
                const int k1 = m.int_ptr[i + 1];
                for (int kb = m.int_ptr[i]; kb < k1; kb += 4) {
                    int inc[4], nb[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        inc[j] = (kb + j < k1) ? m.int_inc[kb + j] : 0;
                        nb[j] = (kb + j < k1) ? m.int_nbr[kb + j] : 0;
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (kb + j >= k1) break;
                        const int e = inc[j] >> 1, side = inc[j] & 1, c2 = nb[j];
                        const double nx = m.enx[e], ny = m.eny[e];
                        const spe::Side so = spe::lr_fwd(h, qx0, qy0, zb0, be0, nx, ny, o.h_min);
                        const spe::Side sn =
                            spe::lr_fwd(U[c2], U[nc + c2], U[2 * nc + c2], m.zb[c2], m.beta[c2], nx, ny, o.h_min);
                        const spe::Side &sL = side ? sn : so, &sR = side ? so : sn;
                        double m1 = fabs(sL.lm), v = fabs(sL.lp);
                        if (m1 < v) m1 = v;
                        if (m1 < (v = fabs(sR.lm))) m1 = v;
                        if (m1 < (v = fabs(sR.lp))) m1 = v;
                        swept += m1 * m.elen[e];
                    }
                }
            }
            for (int k = m.int_ptr[i]; MUSCL && k < m.int_ptr[i + 1]; ++k) {
                const int e = m.int_inc[k] >> 1;
                spe::Side sL, sR;
                e_int_states<MUSCL>(m, o, U, sxy, dt_recon, e, sL, sR);
                double m1 = fabs(sL.lm), v = fabs(sL.lp);
                if (m1 < v) m1 = v;
                if (m1 < (v = fabs(sR.lm))) m1 = v;
                if (m1 < (v = fabs(sR.lp))) m1 = v;
                swept += m1 * m.elen[e];
            }
        }
        const double den = smax_(swept, kSmall);
        const double x = ck->cfl * (m.area[i] / den);
        if (x < cand) cand = x;
    }
    cand = wave_min(cand);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = cand;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < kBlock / 64; ++w) v = (red[w] < v) ? red[w] : v;
        partial[lb] = v;
    }
}

// one block: min of the per-block candidates, then the step schedule
__global__ __launch_bounds__(kDtBlock) void k_e_dt(BcDev bc, EClock *ck, const double *partial,
                                                 int n_part, double *T, double *dtv, double *bcv,
                                                 int *trig, const double *obs_t, int64_t n) {
    if (!(T[n] < ck->t_max)) {  // past t_max: no step, no observation
        if (threadIdx.x == 0) trig[n] = -1;
        return;
    }
    __shared__ double red[kDtBlock / 64];
    const double dt_max = ck->dt_max;
    double v = dt_max;
    for (int k0 = threadIdx.x; k0 < n_part; k0 += 4 * kDtBlock) {  // four loads in flight per thread
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = (k0 + j * kDtBlock < n_part) ? partial[k0 + j * kDtBlock] : dt_max;
#pragma unroll
        for (int j = 0; j < 4; ++j) v = (x[j] < v) ? x[j] : v;
    }
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    // the step's boundary values depend on t alone: one thread per series
    for (int k = (int)threadIdx.x; k < ck->n_series; k += kDtBlock)
        bcv[n * ck->n_series + k] =
            interp_e(bc.t, bc.v + (size_t)k * bc.n_times, bc.n_times, T[n], bc.sorted);
    __syncthreads();
    if (threadIdx.x != 0) return;
    double dtc = red[0];
    for (int w = 1; w < kDtBlock / 64; ++w) dtc = (red[w] < dtc) ? red[w] : dtc;
    const double t = T[n];
    const double t_max = ck->t_max;
    if ((t + dtc) > t_max) dtc = t_max - t;
    if (dtc < 1e-6) dtc = 1e-6;
    double target = dt_max;
    int oi = ck->oi;
    if (oi < ck->n_obs) {
        const double to = obs_t[oi] - t;
        if (to - 10e-6 > 0.0 && to < target) target = to;
    }
    double dt = (dtc < target) ? dtc : target;
    if ((t + dt) > t_max) dt = t_max - t;
    if (dt < 1e-6) dt = 1e-6;
    const double tn = t + dt;
    while (oi < ck->n_obs && tn > obs_t[oi] + 1e-6) oi++;
    int tr = -1;
    if (oi < ck->n_obs && fabs(tn - obs_t[oi]) < 1e-6) tr = oi++;
    ck->oi = oi;
    trig[n] = tr;
    dtv[n] = dt;
    T[n + 1] = tn;
}

struct CellSt {
    double h, qx, qy;
};

// balance of one cell (hypComputeBalance_E, swSpE_hyp.cpp:336-394) before the
// round-off clean-up, with the edge fluxes computed here (hypBoundaryFlux_E,
// hypHllc_E): boundary edges first, then internal edges, each in reference
// order.  Every internal edge's flux is computed by both of its cells, from
// the same inputs by the same operations, so both see the same values.
// bcv: the step's boundary values per series; Fb (optional): the boundary
// fluxes out (each boundary edge belongs to one cell), for the reverse sweep;
// dint (optional): sum of the internal-edge contributions alone, in order.
template <bool MUSCL>
__device__ __forceinline__ CellSt e_balance(const MeshDev &m, const Opts &o, const Controls &cs,
                                            const BcDev &bc, const double *__restrict__ U,
                                            const double *__restrict__ sxy, double dt_recon,
                                            const double *__restrict__ bcv, int i, double dt,
                                            double *__restrict__ Fb, CellSt *dint = nullptr) {
    const int nc = m.n_cell;
    CellSt s{U[i], U[nc + i], U[2 * nc + i]};
    const double phi = cs.phi[i], area = m.area[i];
    for (int k = m.bnd_ptr[i]; k < m.bnd_ptr[i + 1]; ++k) {
        const int j = m.bnd_inc[k];
        const double nx = m.bnx[j], ny = m.bny[j];
        const spe::Side sd = e_bnd_state<MUSCL>(m, o, U, i, j);
        const int ser = m.bser[j];
        double f[3];
        spe::bnd_fwd(o.h_min, o.phi_min, m.beta[i], phi, m.zb[i], sd.z, sd.q, sd.r, sd.lp, bc.type[ser],
                     bcv[ser], f);
        if (Fb) {
            Fb[3 * (size_t)j] = f[0];
            Fb[3 * (size_t)j + 1] = f[1];
            Fb[3 * (size_t)j + 2] = f[2];
        }
        const double co = spe::edge_coeff(m.blen[j], area, phi, o.phi_min);
        s.h += (+co) * f[0] * dt;
        s.qx += ((+co * nx) * f[1] + (+co * nx) * 0.0 + (-co * ny) * f[2]) * dt;
        s.qy += ((+co * ny) * f[1] + (+co * ny) * 0.0 + (+co * nx) * f[2]) * dt;
    }
    if (!MUSCL) {
        // the cell's own data once; per incidence only the edge and the
        // neighbour (int_nbr), the indices of up to 4 incidences loaded
        // together; the same operands and operations as the MUSCL branch
        const double h0 = U[i], qx0 = U[nc + i], qy0 = U[2 * nc + i], zb0 = m.zb[i], be0 = m.beta[i];
        const int k1 = m.int_ptr[i + 1];
        for (int kb = m.int_ptr[i]; kb < k1; kb += 4) {
            int inc[4], nb[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                inc[j] = (kb + j < k1) ? m.int_inc[kb + j] : 0;
                nb[j] = (kb + j < k1) ? m.int_nbr[kb + j] : 0;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (kb + j >= k1) break;
                const int e = inc[j] >> 1, side = inc[j] & 1, c2 = nb[j];
                const double nx = m.enx[e], ny = m.eny[e];
                const double zb2 = m.zb[c2], be2 = m.beta[c2], ph2 = cs.phi[c2];
                const spe::Side so = spe::lr_fwd(h0, qx0, qy0, zb0, be0, nx, ny, o.h_min);
                const spe::Side sn = spe::lr_fwd(U[c2], U[nc + c2], U[2 * nc + c2], zb2, be2, nx, ny, o.h_min);
                const spe::Side &sL = side ? sn : so, &sR = side ? so : sn;
                double f[5];
                spe::hllc_fwd(o.h_min, o.phi_min, side ? be2 : be0, side ? be0 : be2, side ? ph2 : phi,
                              side ? phi : ph2, side ? zb2 : zb0, side ? zb0 : zb2, sL.z, sL.q, sL.r, sR.z,
                              sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, f);
                const double co = cs.coe[inc[j]];  // edge_coeff(elen[e], area, phi)
                double a, b, c;
                spe::side_contrib(side, co, nx, ny, f, dt, a, b, c);
                s.h += a;
                s.qx += b;
                s.qy += c;
                if (dint) {
                    dint->h += a;
                    dint->qx += b;
                    dint->qy += c;
                }
            }
        }
        return s;
    }
    for (int k = m.int_ptr[i]; k < m.int_ptr[i + 1]; ++k) {
        const int inc = m.int_inc[k];
        const int e = inc >> 1, side = inc & 1;
        const int L = m.eL[e], R = m.eR[e];
        spe::Side sL, sR;
        e_int_states<MUSCL>(m, o, U, sxy, dt_recon, e, sL, sR);
        double f[5];
        spe::hllc_fwd(o.h_min, o.phi_min, m.beta[L], m.beta[R], cs.phi[L], cs.phi[R], m.zb[L], m.zb[R],
                      sL.z, sL.q, sL.r, sR.z, sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, f);
        const double co = spe::edge_coeff(m.elen[e], area, phi, o.phi_min);
        double a, b, c;
        spe::side_contrib(side, co, m.enx[e], m.eny[e], f, dt, a, b, c);
        s.h += a;
        s.qx += b;
        s.qy += c;
        if (dint) {
            dint->h += a;
            dint->qx += b;
            dint->qy += c;
        }
    }
    return s;
}

// the same balance from the fluxes k_e_edges_muscl stored (5 per internal
// edge: F1 and the two sides' momentum sums; then 3 per boundary edge):
// side_contrib's operations, its sums taken from the edge pass
__device__ __forceinline__ CellSt e_balance_stored(const MeshDev &m, const Opts &o, const Controls &cs,
                                                   const double *__restrict__ U,
                                                   const double *__restrict__ F, int i, double dt,
                                                   CellSt *dint = nullptr) {
    const int nc = m.n_cell;
    CellSt s{U[i], U[nc + i], U[2 * nc + i]};
    const double phi = cs.phi[i], area = m.area[i];
    const double *Fb = F + 5 * (size_t)m.n_int;
    for (int k = m.bnd_ptr[i]; k < m.bnd_ptr[i + 1]; ++k) {
        const int j = m.bnd_inc[k];
        const double nx = m.bnx[j], ny = m.bny[j];
        const double *f = Fb + 3 * (size_t)j;
        const double co = spe::edge_coeff(m.blen[j], area, phi, o.phi_min);
        s.h += (+co) * f[0] * dt;
        s.qx += ((+co * nx) * f[1] + (+co * nx) * 0.0 + (-co * ny) * f[2]) * dt;
        s.qy += ((+co * ny) * f[1] + (+co * ny) * 0.0 + (+co * nx) * f[2]) * dt;
    }
    // four incidences at a time: their indices, then their edge lengths and
    // stored values, then the sums in incidence order
    const int ie = m.int_ptr[i + 1];
    for (int kb = m.int_ptr[i]; kb < ie; kb += 4) {
        int in[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) in[j] = (kb + j < ie) ? m.int_inc[kb + j] : 0;
        double len[4], f0[4], f1[4], f2[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (kb + j < ie) {
                const int e = in[j] >> 1, side = in[j] & 1;
                const double *f = F + 5 * (size_t)e;
                len[j] = m.elen[e];
                f0[j] = f[0];
                f1[j] = f[1 + 2 * side];
                f2[j] = f[2 + 2 * side];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (kb + j < ie) {
                const double co = spe::edge_coeff(len[j], area, phi, o.phi_min);
                const double a = ((in[j] & 1) ? co : -co) * (f0[j] * dt);
                const double b = f1[j] * dt;
                const double c = f2[j] * dt;
                s.h += a;
                s.qx += b;
                s.qy += c;
                if (dint) {
                    dint->h += a;
                    dint->qx += b;
                    dint->qy += c;
                }
            }
        }
    }
    return s;
}

// per cell: edge fluxes and balance, clean-up, clamp, friction _E, head loss
// _E.  STORED (MUSCL-EVR, dry run and replay): the fluxes come from
// k_e_edges_muscl in Fst instead.  The dry run (T given) takes the boundary
// values the schedule kernel recorded for the step and reconstructs (MUSCL)
// with the previous step's dt (*dtp, 0 before the first); the replay takes
// dtr, the step's own.  STORE (a replay before a reverse sweep, or a
// Godunov dry run that keeps the resident trajectory): also keep the
// balanced state before the clean-up and the sum of its internal-edge
// increments, P = [pre | dint], and (computing the fluxes here) the
// boundary fluxes Fb.
// (the Godunov dry run that keeps the resident trajectory takes 125 VGPRs,
// 4 waves per SIMD: held at 96 for 5 waves, its incidence loop spills)
template <bool STORE, bool MUSCL, bool STORED = false>
__global__ __launch_bounds__(kBlock) void k_e_update(MeshDev m, Opts o, Controls cs, BcDev bc,
                                                     const double *__restrict__ Uin,
                                                     double *__restrict__ Uout,
                                                     const double *__restrict__ sxy, const double *dtp,
                                                     double dtr, const double *__restrict__ bcv,
                                                     const double *dtv, double *T, int64_t n,
                                                     double t_max, double *__restrict__ P,
                                                     double *__restrict__ Fb,
                                                     const double *__restrict__ Fst,
                                                     GaugeSample gs) {
    const int nc = m.n_cell;
    if (T && !(T[n] < t_max)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) T[n + 1] = T[n];
        const int i = xcd_block() * kBlock + (int)threadIdx.x;
        if (i < nc) {
            Uout[i] = Uin[i];
            Uout[nc + i] = Uin[nc + i];
            Uout[2 * nc + i] = Uin[2 * nc + i];
        }
        return;
    }
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= nc) return;
    const double dt = dtv[n];
    const double dt_recon = MUSCL ? (T ? (dtp ? *dtp : 0.0) : dtr) : 0.0;
    CellSt dint{0.0, 0.0, 0.0};
    CellSt s = STORED ? e_balance_stored(m, o, cs, Uin, Fst, i, dt, STORE ? &dint : nullptr)
                      : e_balance<MUSCL>(m, o, cs, bc, Uin, sxy, dt_recon, bcv, i, dt, Fb,
                                         STORE ? &dint : nullptr);
    if (STORE) {
        P[i] = s.h;
        P[nc + i] = s.qx;
        P[2 * nc + i] = s.qy;
        P[3 * nc + i] = dint.h;
        P[4 * nc + i] = dint.qx;
        P[5 * nc + i] = dint.qy;
    }
    if (fabs(s.h) < kSmall) s.h = 0.0;
    if (fabs(s.qx) < kSmall) s.qx = 0.0;
    if (fabs(s.qy) < kSmall) s.qy = 0.0;
    if (s.h < 0.) {
        s.h = 0.;
        s.qx = 0.;
        s.qy = 0.;
    }
    if (o.friction != sp::FRIC_NONE)
        spe::fric_fwd(o.h_min, s.h, s.qx, s.qy, m.fcos[i], m.fsin[i], cs.K1[i], cs.K2[i], dt);
    if (o.head_loss)
        spe::hl_fwd(o.h_min, s.h, s.qx, s.qy, m.hcos[i], m.hsin[i], m.hsing[i], m.hk1[i], m.hk2[i],
                    dt);
    Uout[i] = s.h;
    Uout[nc + i] = s.qx;
    Uout[2 * nc + i] = s.qy;
    if (gs.gcell) {
        const int oi = gs.trig[n];
        if (oi >= 0)
            for (int g = 0; g < gs.n_g; ++g)
                if (gs.gcell[g] == i) gs.gser[(size_t)oi * gs.n_g + g] = s.h;
    }
}

// gauge cost (compute_cost_gauges_ptr, swSpE_gauges.cpp:166-179): one thread,
// fixed order.  Jo[0] += sum_g w_g (h - obs)^2 / norm (Jo = slot of the observation)
__global__ void k_cost_gauges(const double *__restrict__ U, const int *gcell, const double *gw,
                              const double *gobs, int n_g, double norm, double *J) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double Jo = 0.0;
    for (int g = 0; g < n_g; ++g) {
        const double d = U[gcell[g]] - gobs[g];
        Jo += gw[g] * d * d;
    }
    J[0] += Jo / norm;
}

// total cost: observation costs summed in observation (= time) order, so the
// result does not depend on the order segments were replayed in
__global__ void k_cost_total(double *J, int n_obs) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < n_obs; ++k) s += J[1 + k];
    J[0] = s;
}

// frame cost (compute_cost_frame_ptr, swSpE_ad.cpp:273-308): per-block partial
// sums, then one block sums them in block order (deterministic).
__global__ __launch_bounds__(kBlock) void k_cost_frame_part(const double *__restrict__ U, int nc,
                                                             const double *fh, const double *fqx,
                                                             const double *fqy, const double *wh,
                                                             const double *wq, double *part) {
    __shared__ double red[kBlock];
    const int i = blockIdx.x * kBlock + (int)threadIdx.x;
    double v = 0.0;
    if (i < nc) {
        const double dh = U[i] - fh[i], dqx = U[nc + i] - fqx[i], dqy = U[2 * nc + i] - fqy[i];
        v = wh[i] * (dh * dh) + wq[i] * ((dqx * dqx) + (dqy * dqy));
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ||x||^2 of n values: per-block tree sums, then one thread sums the blocks
// in block order (deterministic)
__global__ __launch_bounds__(kBlock) void k_sq_part(const double *__restrict__ x, int n,
                                                    double *part) {
    __shared__ double red[kBlock];
    const int i = blockIdx.x * kBlock + (int)threadIdx.x;
    const double v = (i < n) ? x[i] : 0.0;
    red[threadIdx.x] = v * v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// one block of kBlock threads: thread t sums blocks t, t + kBlock, ... in
// order, then a fixed tree (deterministic)
__global__ __launch_bounds__(kBlock) void k_sq_sum(const double *part, int nparts, double *out) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int b = (int)threadIdx.x; b < nparts; b += kBlock) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_cost_frame_sum(const double *part, int nparts, double norm, double *J) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < nparts; ++k) s += part[k];
    J[0] += s / norm;
}

// full state after a step that reaches an observation time (the frame the
// reference's storeReferenceMap writes at that trigger)
__global__ void k_frame_sample(const double *__restrict__ U, int64_t nc, const int *trig, int64_t n,
                               double *frames) {
    const int oi = trig[n];
    if (oi < 0) return;
    double *dst = frames + (size_t)oi * 3 * nc;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < 3 * nc;
         q += (int64_t)gridDim.x * blockDim.x)
        dst[q] = U[q];
}

// ---------------------------------------------------------------------------
// reverse kernels
// ---------------------------------------------------------------------------

// cost seed of one observation, applied where the adjoint of the observed
// state is formed: dJ_obs/dU of the gauge misfit (cells chained through
// gfirst/gnext in gauge order) and of the frame misfit
struct SeedArgs {
    int oi;                               // observation index, -1 = none
    const double *U;                      // the observed state
    const int *gfirst, *gnext;            // per-cell gauge chain (-1 ends)
    const double *gw, *gobs;              // weights, this observation's depths
    const double *fh, *fqx, *fqy, *wh, *wq;  // this observation's frame (fh = nullptr: none)
    double norm;
};

__device__ __forceinline__ void seed_cell(const SeedArgs &sd, int nc, int i, double &ah, double &ax,
                                          double &ay) {
    if (sd.oi < 0) return;
    if (sd.gfirst)
        for (int g = sd.gfirst[i]; g >= 0; g = sd.gnext[g]) ah += 2.0 * sd.gw[g] * (sd.U[i] - sd.gobs[g]) / sd.norm;
    if (sd.fh) {
        ah += 2.0 * sd.wh[i] * (sd.U[i] - sd.fh[i]) / sd.norm;
        ax += 2.0 * sd.wq[i] * (sd.U[nc + i] - sd.fqx[i]) / sd.norm;
        ay += 2.0 * sd.wq[i] * (sd.U[2 * nc + i] - sd.fqy[i]) / sd.norm;
    }
}

// Reverse of the cell-local part of one step (sources, clamp, clean-up) and
// the porosity derivative of its internal-edge factors.  (ah, ax, ay) in =
// adjoint of U^{n+1}; out: lams = adjoint of the balanced state, which is
// also the cell-local part of the adjoint of U^n (boundary cells: completed
// into part by the boundary blocks of k_r_edge).  P = [balanced state |
// internal-edge increments] from the replay.
__device__ __forceinline__ void cell_reverse(const MeshDev &m, const Opts &o, const Controls &cs,
                                             const double *__restrict__ P, double dt, int i,
                                             double ah, double ax, double ay,
                                             double *__restrict__ lams, double &gp, double *gK1,
                                             double *gK2) {
    const int nc = m.n_cell;
    const CellSt dint{P[3 * nc + i], P[4 * nc + i], P[5 * nc + i]};
    CellSt s{P[i], P[nc + i], P[2 * nc + i]};
    const bool kh = !(fabs(s.h) < kSmall), kx = !(fabs(s.qx) < kSmall), ky = !(fabs(s.qy) < kSmall);
    if (!kh) s.h = 0.0;
    if (!kx) s.qx = 0.0;
    if (!ky) s.qy = 0.0;
    const bool clamped = s.h < 0.;
    if (clamped) {
        s.h = 0.;
        s.qx = 0.;
        s.qy = 0.;
    }
    const CellSt after_clamp = s;
    if (s.h < o.adj_cut) {  // thin film: its adjoint is not carried back (adj_cut = 0: never)
        ah = 0.0;
        ax = 0.0;
        ay = 0.0;
    }
    if (o.head_loss) {
        if (o.friction != sp::FRIC_NONE)
            spe::fric_fwd(o.h_min, s.h, s.qx, s.qy, m.fcos[i], m.fsin[i], cs.K1[i], cs.K2[i], dt);
        spe::hl_vjp(o.h_min, s.h, s.qx, s.qy, m.hcos[i], m.hsin[i], m.hsing[i], m.hk1[i], m.hk2[i],
                    dt, ah, ax, ay);
    }
    if (o.friction != sp::FRIC_NONE) {
        double dk1 = 0.0, dk2 = 0.0;
        spe::fric_vjp(o.h_min, after_clamp.h, after_clamp.qx, after_clamp.qy, m.fcos[i], m.fsin[i],
                      cs.K1[i], cs.K2[i], dt, ah, ax, ay, dk1, dk2);
        gK1[i] += dk1;
        gK2[i] += dk2;
    }
    if (clamped) {
        ah = 0.0;
        ax = 0.0;
        ay = 0.0;
    }
    if (!kh) ah = 0.0;
    if (!kx) ax = 0.0;
    if (!ky) ay = 0.0;
    lams[i] = ah;
    lams[nc + i] = ax;
    lams[2 * nc + i] = ay;
    // balance: identity part; internal-edge factors co_e = w_e / area / phi
    // enter every contribution linearly, so
    // sum_e (dJ/dco_e) (dco_e/dphi) = -(lam . dint) / phi
    const double phi = cs.phi[i];
    if (phi > o.phi_min) gp += -((ah * dint.h + ax * dint.qx) + ay * dint.qy) / phi;
}

// One fused reverse kernel per step boundary n:
//   gather  the adjoint of U^n = the cell part (lams, or part for a cell
//           with boundary edges) + the edge-side contributions of step
//           n in the cell's reference incidence order (reset: the tile top,
//           the adjoint restarts from 0; the porosity terms still count)
//   seed    + dJ_obs/dU^n when U^n is observed
//   CELL    reverse of the cell part of step n-1 (else: write lam = adj U^n)
// gather: -1 nothing to gather (the last step, or a tile top: the future
// adjoint is dropped), 1 gather the cell part + edge sides (+ zn, MUSCL-EVR:
// the neighbours' slope adjoints through this cell's free surface), 2 take lam
// (the gather already done by the CELL = false launch that closed the next
// segment).
// Blocks past n_cell_blocks sum the boundary-value adjoints of step n per
// series (one block per series, fixed order).
template <bool CELL>
__global__ __launch_bounds__(kBlock) void k_r_gc(MeshDev m, Opts o, Controls cs,
                                                 double *__restrict__ part,
                                                 const double *__restrict__ side, int gather,
                                                 SeedArgs sd, const double *__restrict__ P,
                                                 double dt, double *__restrict__ lams,
                                                 double *__restrict__ lam, double *gphi,
                                                 double *gK1, double *gK2, int n_cell_blocks,
                                                 const int *ser_ptr, const int *ser_edges,
                                                 const double *gbce, double *gbc_n,
                                                 const double *__restrict__ zn) {
    if ((int)blockIdx.x >= n_cell_blocks) {
        __shared__ double red[kBlock];
        const int k = (int)blockIdx.x - n_cell_blocks;
        double v = 0.0;
        for (int q = ser_ptr[k] + (int)threadIdx.x; q < ser_ptr[k + 1]; q += kBlock) v += gbce[ser_edges[q]];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) gbc_n[k] = red[0];
        return;
    }
    const int i = xcd_remap((int)blockIdx.x, n_cell_blocks) * kBlock + (int)threadIdx.x;
    const int nc = m.n_cell;
    if (i >= nc) return;
    double ah = 0.0, ax = 0.0, ay = 0.0, gp = 0.0;
    if (gather == 2) {  // gathered at the end of the previous segment
        ah = lam[i];
        ax = lam[nc + i];
        ay = lam[2 * nc + i];
    } else if (gather >= 0) {
        // the cell part: lams as the cell reverse left it, or, for a cell
        // with boundary edges, part (lams plus its boundary terms)
        const bool bcell = m.bnd_ptr[i] < m.bnd_ptr[i + 1];
        const double a0 = bcell ? part[i] : lams[i];
        const double b0 = bcell ? part[nc + i] : lams[nc + i];
        const double c0 = bcell ? part[2 * nc + i] : lams[2 * nc + i];
        double a = a0, b = b0, c = c0;
        // four incidences at a time: their indices, then their side records
        // (and slope terms), then the sums in incidence order
        const int ie = m.int_ptr[i + 1];
        for (int kb = m.int_ptr[i]; kb < ie; kb += 4) {
            int in[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) in[j] = (kb + j < ie) ? m.int_inc[kb + j] : 0;
            double s0[4], s1[4], s2[4], s3[4], z[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (kb + j < ie) {
                    const double *sv = side + 8 * (size_t)(in[j] >> 1) + 4 * (in[j] & 1);
                    s0[j] = sv[0];
                    s1[j] = sv[1];
                    s2[j] = sv[2];
                    s3[j] = sv[3];
                    z[j] = zn ? zn[in[j]] : 0.0;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (kb + j < ie) {
                    a += s0[j];
                    b += s1[j];
                    c += s2[j];
                    gp += s3[j];
                    if (zn) a += z[j];  // the neighbour's slope through this cell's surface
                }
            }
        }
        if (gather > 0) {  // not a tile top
            ah = a;
            ax = b;
            ay = c;
        }
    }
    seed_cell(sd, nc, i, ah, ax, ay);
    // the gathered porosity terms and those of the cell reverse are added
    // one after the other, so the result does not depend on whether a
    // segment boundary splits them over two launches (stride invariance)
    double g = gphi[i] + gp;
    if (CELL) {
        double gc_ = 0.0;
        cell_reverse(m, o, cs, P, dt, i, ah, ax, ay, lams, gc_, gK1, gK2);
        g += gc_;
    } else {
        lam[i] = ah;
        lam[nc + i] = ax;
        lam[2 * nc + i] = ay;
    }
    gphi[i] = g;
}

// the boundary part of the edge reverse of a step: per boundary cell, the
// reverse of its boundary fluxes (hypBoundaryFlux_E and the Godunov or
// MUSCL-EVR _E boundary states) and of their factors, in reference order,
// added to the cell part of the adjoint (lams) into part; boundary-value
// adjoint per edge
struct BndRev {
    BcDev bc;
    const double *F;    // the step's boundary fluxes, 3 per boundary edge
    const double *bcv;  // the step's boundary values per series
    double *part, *gphi, *gbce;
    int n_bnd_blocks;   // leading blocks of k_r_edge that run it (a multiple of 8)
};

template <bool MUSCL>
__device__ __forceinline__ void bnd_reverse(const MeshDev &m, const Opts &o, const Controls &cs,
                                            const BndRev &br, const double *__restrict__ U, double dt,
                                            const double *__restrict__ lams, int t) {
    if (t >= m.n_bnd_cells) return;
    const BcDev &bc = br.bc;
    const double *F = br.F, *bcv = br.bcv;
    double *part = br.part, *gphi = br.gphi, *gbce = br.gbce;
    const int i = m.bnd_cells[t];
    const int nc = m.n_cell;
    const double ah = lams[i], ax = lams[nc + i], ay = lams[2 * nc + i];
    double ph = ah, px = ax, py = ay, gp = 0.0;  // the cell part (lams), completed here
    const double phi = cs.phi[i], area = m.area[i];
    const double h0 = U[i], qx0 = U[nc + i], qy0 = U[2 * nc + i];
    const double *Fb = F;  // the step's boundary fluxes, 3 per boundary edge
    for (int k = m.bnd_ptr[i]; k < m.bnd_ptr[i + 1]; ++k) {
        const int j = m.bnd_inc[k];
        const double nx = m.bnx[j], ny = m.bny[j];
        const double *f = Fb + 3 * (size_t)j;
        const spe::Side sd = MUSCL ? spe::lr_bnd_muscl(h0, qx0, qy0, m.zb[i], m.beta[i], nx, ny)
                                   : spe::lr_fwd(h0, qx0, qy0, m.zb[i], m.beta[i], nx, ny, o.h_min);
        const int ser = m.bser[j];
        const double co = spe::edge_coeff(m.blen[j], area, phi, o.phi_min);
        double df[3];
        df[0] = ah * co * dt;
        df[1] = dt * (ax * co * nx + ay * co * ny);
        df[2] = dt * (-ax * co * ny + ay * co * nx);
        const double dco = ah * f[0] * dt + dt * (ax * (nx * f[1] - ny * f[2]) + ay * (ny * f[1] + nx * f[2]));
        gp += dco * spe::edge_coeff_dphi(m.blen[j], area, phi, o.phi_min);
        double d[6] = {0, 0, 0, 0, 0, 0};
        spe::bnd_vjp(o.h_min, o.phi_min, m.beta[i], phi, m.zb[i], sd.z, sd.q, sd.r, sd.lp,
                     bc.type[ser], bcv[ser], df, d);
        if (MUSCL)
            spe::lr_bnd_muscl_vjp(h0, qx0, qy0, m.beta[i], nx, ny, d[0], d[1], d[2], d[3], ph, px, py);
        else
            spe::lr_vjp(h0, qx0, qy0, m.beta[i], nx, ny, o.h_min, d[0], d[1], d[2], 0.0, d[3], ph, px, py);
        gp += d[4];
        gbce[j] = d[5];
    }
    part[i] = ph;
    part[nc + i] = px;
    part[2 * nc + i] = py;
    gphi[i] += gp;
}

// an opaque copy of x: expressions of the copy are not merged with the same
// expressions of x evaluated earlier (an empty asm statement, no instruction)
__device__ __forceinline__ double fresh(double x) {
    asm volatile("" : "+v"(x));
    return x;
}

// per internal edge: flux adjoints from both sides, HLLC_E reverse, then the
// reverse of the Godunov _E states, or (MUSCL) of the MUSCL-EVR
// reconstruction of each side from the step's slopes sxy with the step's dt,
// whose slope adjoints go to sd[4e ..] = {L: dsx, dsy, R: dsx, dsy}
// register budgets: 4 waves per SIMD (128 VGPRs) fit the Godunov reverse
// without spills, the MUSCL-EVR one needs 144 (3 waves).  The first
// br.n_bnd_blocks blocks run the boundary reverse of the step, dispatched
// first so that its latency overlaps the edges (a few blocks: the spills its
// larger register need takes under this budget stay there); a multiple of 8,
// so that the edge blocks keep their XCD-aware order
template <bool MUSCL>
__global__ __launch_bounds__(kBlock, MUSCL ? 3 : 4) void k_r_edge(MeshDev m, Opts o, Controls cs,
                                                   const double *__restrict__ U, double dt,
                                                   const double *__restrict__ lams,
                                                   double *__restrict__ side_out,
                                                   const double *__restrict__ sxy,
                                                   double *__restrict__ sd, BndRev br) {
    if ((int)blockIdx.x < br.n_bnd_blocks) {
        bnd_reverse<MUSCL>(m, o, cs, br, U, dt, lams, (int)blockIdx.x * kBlock + (int)threadIdx.x);
        return;
    }
    const int e = xcd_remap((int)blockIdx.x - br.n_bnd_blocks, (int)gridDim.x - br.n_bnd_blocks) * kBlock +
                  (int)threadIdx.x;
    if (e >= m.n_int) return;
    const int nc = m.n_cell;
    const int L = m.eL[e], R = m.eR[e];
    const double nx = m.enx[e], ny = m.eny[e];
    const double phiL = cs.phi[L], phiR = cs.phi[R];
    const double coL = cs.coe[2 * (size_t)e], coR = cs.coe[2 * (size_t)e + 1];
    const double Fz[5] = {0, 0, 0, 0, 0};
    double dF[5] = {0, 0, 0, 0, 0};
    spe::side_contrib_vjp(0, coL, nx, ny, Fz, dt, lams[L], lams[nc + L], lams[2 * nc + L], dF);
    spe::side_contrib_vjp(1, coR, nx, ny, Fz, dt, lams[R], lams[nc + R], lams[2 * nc + R], dF);
    const double hL = U[L], qxL = U[nc + L], qyL = U[2 * nc + L];
    const double hR = U[R], qxR = U[nc + R], qyR = U[2 * nc + R];
    const double bL = m.beta[L], bR = m.beta[R];
    double d[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dhL = 0, dqxL = 0, dqyL = 0, dhR = 0, dqxR = 0, dqyR = 0;
    if (MUSCL) {
        const double xm = m.exm[e], ym = m.eym[e];
        const double sxL = sxy[L], syL = sxy[nc + L], sxR = sxy[R], syR = sxy[nc + R];
        const double dxL = xm - m.cx[L], dyL = ym - m.cy[L], dxR = xm - m.cx[R], dyR = ym - m.cy[R];
        const spe::Side sL = spe::recon_muscl(0, hL, qxL, qyL, m.zb[L], bL, sxL, syL, nx, ny, dxL, dyL, dt);
        const spe::Side sR = spe::recon_muscl(1, hR, qxR, qyR, m.zb[R], bR, sxR, syR, nx, ny, dxR, dyR, dt);
        spe::hllc_vjp(o.h_min, o.phi_min, bL, bR, phiL, phiR, m.zb[L], m.zb[R], sL.z, sL.q, sL.r, sR.z,
                      sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, dF, d);
        double gxL = 0, gyL = 0, gxR = 0, gyR = 0;
        // the VJPs recompute the reconstruction from fresh copies of its
        // inputs instead of keeping its intermediates live across hllc_vjp
        // (216 -> 144 VGPRs, 2 -> 3 waves per SIMD)
        spe::recon_muscl_vjp(0, fresh(hL), fresh(qxL), fresh(qyL), fresh(m.zb[L]), bL, fresh(sxL),
                             fresh(syL), nx, ny, fresh(dxL), fresh(dyL), dt, d[0], d[1], d[2], d[6],
                             d[8], dhL, dqxL, dqyL, gxL, gyL);
        spe::recon_muscl_vjp(1, fresh(hR), fresh(qxR), fresh(qyR), fresh(m.zb[R]), bR, fresh(sxR),
                             fresh(syR), nx, ny, fresh(dxR), fresh(dyR), dt, d[3], d[4], d[5], d[7],
                             d[9], dhR, dqxR, dqyR, gxR, gyR);
        double *g = sd + 4 * (size_t)e;
        g[0] = gxL;
        g[1] = gyL;
        g[2] = gxR;
        g[3] = gyR;
    } else {
        spe::LrAux xL, xR;
        const spe::Side sL = spe::lr_fwd_x(hL, qxL, qyL, m.zb[L], bL, nx, ny, o.h_min, xL);
        const spe::Side sR = spe::lr_fwd_x(hR, qxR, qyR, m.zb[R], bR, nx, ny, o.h_min, xR);
        spe::hllc_vjp(o.h_min, o.phi_min, bL, bR, phiL, phiR, m.zb[L], m.zb[R], sL.z, sL.q, sL.r, sR.z,
                      sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, dF, d);
        spe::lr_vjp_x(xL, bL, nx, ny, d[0], d[1], d[2], d[6], d[8], dhL, dqxL, dqyL);
        spe::lr_vjp_x(xR, bR, nx, ny, d[3], d[4], d[5], d[7], d[9], dhR, dqxR, dqyR);
    }
    double *so = side_out + 8 * (size_t)e;
    so[0] = dhL;
    so[1] = dqxL;
    so[2] = dqyL;
    so[3] = d[10];
    so[4] = dhR;
    so[5] = dqxR;
    so[6] = dqyR;
    so[7] = d[11];
}

// the slope adjoint's term for the neighbour across a cell's incidence k
// (edge e) goes to zn[2 e + the neighbour's side], the slot of the
// neighbour's own incidence on e: each slot has one writer, and the gather
// reads it at the index of its incidence, next to the edge's side record
struct ZnOut {
    double *zn;
    const int *inc;  // the cell's incidences: entry (e << 1) | side
    __device__ size_t slot(int k) const { return (size_t)(inc[k] ^ 1); }
    __device__ void set(int k, double v) const { zn[slot(k)] = v; }
    __device__ void add(int k, double v) const { zn[slot(k)] += v; }
    __device__ void put(int k, double v) const { zn[slot(k)] = v; }
};

// cells of at most this many internal edges (every triangle or quad) keep
// their incidences in registers in the slope reverse
constexpr int kSlopeRegDeg = 4;

// MUSCL-EVR, per cell: gather the adjoint of its limited slope from its
// edges' sides (sd), reverse the slope and its limiter (spe::slope_vjp): the
// term of its own free surface is added to the h part of the adjoint of U^n,
// those of its neighbours go to zn (CSR slots of this cell)
__global__ __launch_bounds__(kBlock) void k_r_slopes(MeshDev m, const double *__restrict__ U,
                                                     const double *__restrict__ sd,
                                                     double *__restrict__ part,
                                                     double *__restrict__ lams,
                                                     double *__restrict__ zn) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= m.n_cell) return;
    const int b = m.int_ptr[i], deg = m.int_ptr[i + 1] - b;
    const ZnOut out{zn, m.int_inc + b};
    const double zi = U[i] + m.zb[i], xi = m.cx[i], yi = m.cy[i];
    double gx = 0.0, gy = 0.0, dzi = 0.0;
    if (deg <= kSlopeRegDeg) {
        // every incidence's slope adjoints, neighbour and edge data loaded
        // once, together; the slope adjoints summed in incidence order
        static_assert(kSlopeRegDeg == 4, "NbReg holds 4 incidences");
        NbReg r;
        int in[4], nb[4];
        double g0[4], g1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            in[j] = (j < deg) ? m.int_inc[b + j] : 0;
            nb[j] = (j < deg) ? m.int_nbr[b + j] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j < deg) {
                const int c2 = nb[j], e = in[j] >> 1;
                g0[j] = sd[2 * (size_t)in[j]];
                g1[j] = sd[2 * (size_t)in[j] + 1];
                r.z[j] = U[c2] + m.zb[c2];
                r.x[j] = m.cx[c2];
                r.y[j] = m.cy[c2];
                r.xm_[j] = m.exm[e];
                r.ym_[j] = m.eym[e];
                r.s[j] = in[j] & 1;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j < deg) {
                gx += g0[j];
                gy += g1[j];
            }
        }
        spe::slope_vjp_reg<kSlopeRegDeg>(r, deg, zi, xi, yi, gx, gy, dzi, out);
    } else {
        for (int k = 0; k < deg; ++k) {
            const int inc = m.int_inc[b + k];
            gx += sd[2 * (size_t)inc];
            gy += sd[2 * (size_t)inc + 1];
        }
        const NbDev nb{m, U, b};
        spe::slope_vjp(nb, deg, zi, xi, yi, gx, gy, dzi, out);
    }
    // into the cell part of the adjoint of U^n: part for a cell with
    // boundary edges, lams otherwise (the edge reverse has read lams)
    if (m.bnd_ptr[i] < m.bnd_ptr[i + 1])
        part[i] += dzi;
    else
        lams[i] += dzi;
}

// the edge-side factors of the current phi (Controls::coe)
__global__ __launch_bounds__(kBlock) void k_coe(MeshDev m, Opts o, const double *__restrict__ phi,
                                                double *__restrict__ coe) {
    const int e = (int)blockIdx.x * kBlock + (int)threadIdx.x;
    if (e >= m.n_int) return;
    const int L = m.eL[e], R = m.eR[e];
    const double len = m.elen[e];
    coe[2 * (size_t)e] = spe::edge_coeff(len, m.area[L], phi[L], o.phi_min);
    coe[2 * (size_t)e + 1] = spe::edge_coeff(len, m.area[R], phi[R], o.phi_min);
}

__global__ void k_zero(double *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
int zero(sw2dg_ctx *c, double *p, size_t n) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, p, n);
    return SW2DG_OK;
}

double phi_correct(int mode, double lam, double lo, double scale, double m) {
    switch (mode) {
        case 1: return (m < 0.0) ? 0.0 : ((m > 1.05) ? 1.05 : m);
        case 2: return 1.0 / (1.0 + std::exp(-lam * m));
        case 3: return (m - lo) * scale;
        default: return m;
    }
}

double dphi_dm(int mode, double lam, double scale, double m) {
    switch (mode) {
        case 1: return (m < 0.0 || m > 1.05) ? 0.0 : 1.0;
        case 2: {
            const double s = 1.0 / (1.0 + std::exp(-lam * m));
            return lam * s * (1.0 - s);
        }
        case 3: return scale;
        default: return 1.0;
    }
}

// lo / hi of m over the free cells (all cells when none are listed); a
// (nearly) uniform map disables the normalisation (phi = m)
void update_norm_bounds(AssimState *a, const double *m, int64_t nc) {
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    auto see = [&](int64_t i) {
        lo = std::min(lo, m[i]);
        hi = std::max(hi, m[i]);
    };
    if (!a->free_idx.empty())
        for (int64_t i : a->free_idx) see(i);
    else
        for (int64_t i = 0; i < nc; ++i) see(i);
    if (hi - lo < 1e-12) {
        a->norm_lo = 0.0;
        a->norm_scale = 1.0;
    } else {
        a->norm_lo = lo;
        a->norm_scale = 1.0 / (hi - lo);
    }
}

int upload_ctl(sw2dg_ctx *c, AssimState *a, const double *phi, const double *K1, const double *K2,
               bool new_bounds = false) {
    const int64_t nc = c->nc;
    std::vector<double> phi_t;
    if (phi && a->phi_mode != 0) {  // phi is the control m: apply phi_correct on the host
        a->m_last.assign(phi, phi + nc);
        if (a->phi_mode == 3 && new_bounds) update_norm_bounds(a, phi, nc);
        phi_t.resize(nc);
        for (int64_t i = 0; i < nc; ++i)
            phi_t[i] = phi_correct(a->phi_mode, a->phi_lambda, a->norm_lo, a->norm_scale, phi[i]);
        phi = phi_t.data();
    } else {
        a->m_last.clear();
    }
    auto put = [&](const double *src, double *dst, const double *dflt) -> int {
        if (src) {
            std::vector<double> v = c->to_dev_cells(src);
            HIPCHK(c, hipMemcpyAsync(dst, v.data(), nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, hipMemcpyAsync(dst, dflt, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        return SW2DG_OK;
    };
    int st;
    if ((st = put(phi, a->d_phi, c->m.phi))) return st;
    if (c->ni > 0)
        KLAUNCH(a, k_coe, dim3(nblocks(c->ni)), dim3(kBlock), 0, c->stream, c->m, c->o, a->d_phi, a->d_coe);
    if ((st = put(K1, a->d_K1, c->m.fc1))) return st;
    if ((st = put(K2, a->d_K2, c->m.fc2))) return st;
    return SW2DG_OK;
}

Controls ctl(AssimState *a) { return Controls{a->d_phi, a->d_K1, a->d_K2, a->d_coe}; }

// boundary fluxes of one step (3 per boundary edge), kept per replayed step
// for the boundary reverse
size_t flux_len(const sw2dg_ctx *c) { return 3 * (size_t)std::max<int64_t>(c->nb, 1); }

// the three parts of a resident block (doubles)
size_t blk_states(const AssimState *a, const sw2dg_ctx *c) { return (size_t)(a->stride + 1) * 3 * c->nc; }
size_t blk_pre(const AssimState *a, const sw2dg_ctx *c) { return (size_t)a->stride * 6 * c->nc; }
size_t blk_len(const AssimState *a, const sw2dg_ctx *c) {
    return blk_states(a, c) + blk_pre(a, c) + (size_t)a->stride * flux_len(c);
}

double *seg_state(AssimState *a, sw2dg_ctx *c, int j) {
    if (a->seg_blk >= 0) return a->res_blk[a->seg_blk] + (size_t)j * 3 * c->nc;
    return a->d_seg + (size_t)j * 3 * c->nc;
}
double *ckpt_state(AssimState *a, sw2dg_ctx *c, int64_t k) {
    return a->d_ckpt + (size_t)k * 3 * c->nc;
}
double *seg_flux(AssimState *a, sw2dg_ctx *c, int j) {
    if (a->seg_blk >= 0)
        return a->res_blk[a->seg_blk] + blk_states(a, c) + blk_pre(a, c) + (size_t)j * flux_len(c);
    return a->d_Fseg + (size_t)j * flux_len(c);
}
double *seg_pre(AssimState *a, sw2dg_ctx *c, int j) {
    if (a->seg_blk >= 0) return a->res_blk[a->seg_blk] + blk_states(a, c) + (size_t)j * 6 * c->nc;
    return a->d_Pseg + (size_t)j * 6 * c->nc;
}
double *seg_slopes(AssimState *a, sw2dg_ctx *c, int j) { return a->d_Sseg + (size_t)j * 2 * c->nc; }

// segment k's resident block, allocated on first use within the budget
// (Godunov only: its dry run and replay are the same computation, while the
// MUSCL-EVR replay reconstructs with another dt); nullptr = replay it
double *resident_block(sw2dg_ctx *c, AssimState *a, int64_t k) {
    if (a->resident_mode == 0 || c->o.scheme == 1) return nullptr;
    if ((int64_t)a->res_blk.size() <= k) a->res_blk.resize(k + 1, nullptr);
    if (a->res_blk[k]) return a->res_blk[k];
    if (a->res_budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return nullptr;
        a->res_budget = (size_t)(0.6 * (double)free_b);
    }
    const size_t b = blk_len(a, c) * sizeof(double);
    if (a->res_bytes + b > a->res_budget) return nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, b) != hipSuccess) {
        (void)hipGetLastError();
        a->res_budget = a->res_bytes;  // no more blocks
        return nullptr;
    }
    a->allocs.push_back(p);
    a->res_bytes += b;
    a->res_blk[k] = (double *)p;
    return a->res_blk[k];
}

// one _E step U_in -> U_out.  Dry run: (MUSCL slopes,) CFL partials, the
// schedule (dt, boundary values, observation trigger), the update.  Replay:
// (MUSCL slopes,) the update on the recorded schedule; Fb (a slot of the
// segment's boundary-flux buffer), P and S keep what the reverse sweep reads.
// The dry run reconstructs (MUSCL) with the previous step's dt (0 before the
// first step), the replay with the step's own.
void launch_e_step(sw2dg_ctx *c, AssimState *a, const double *Uin, double *Uout, int64_t n,
                   bool dry, double *Fb, double *P = nullptr, double *S = nullptr) {
    const dim3 blk(kBlock), gc(nblocks(c->nc));
    const Controls cs = ctl(a);
    const int ns = std::max(c->n_series, 1);
    const bool muscl = c->o.scheme == 1;
    const double *bcv = a->d_bc + (size_t)n * ns;
    double *sxy = S ? S : a->d_sxy;
    const double *dtp = (dry && n > 0) ? a->d_dt + (n - 1) : nullptr;
    const double dtr = dry ? 0.0 : a->h_dt[n];
    double *T = dry ? a->d_T : nullptr;
    if (muscl)
        KLAUNCH(a, k_e_slopes, gc, blk, 0, c->stream, c->m, Uin, sxy, T, n, c->prm.t_max);
    if (dry) {
        if (muscl) {  // edges once, their fluxes and swept terms kept
            KLAUNCH(a, k_e_edges_muscl, dim3(nblocks(c->ni + c->nb)), blk, 0, c->stream, c->m, c->o, cs,
                    c->bc, Uin, a->d_F, a->d_sw, a->d_T, n, c->prm.t_max, sxy, dtp, 0.0, nullptr, nullptr);
            KLAUNCH(a, k_e_cfl_sw, gc, blk, 0, c->stream, c->m, c->o, Uin, a->d_sw, a->d_ck, a->d_partial,
                    a->d_T, n);
        } else {
            KLAUNCH(a, k_e_cfl<false>, gc, blk, 0, c->stream, c->m, c->o, Uin, sxy, dtp, a->d_ck,
                    a->d_partial, a->d_T, n);
        }
        KLAUNCH(a, k_e_dt, dim3(1), dim3(kDtBlock), 0, c->stream, c->bc, a->d_ck, a->d_partial, nblocks(c->nc),
                a->d_T, a->d_dt, a->d_bc, a->d_trig, a->d_obs_t, n);
    }
    GaugeSample gs{};
    if (dry && a->n_gauges > 0) gs = GaugeSample{a->d_trig, a->d_gcell, a->n_gauges, a->d_gser};
#define E_UPDATE(STORE, MUS)                                                                       \
    KLAUNCH_NAMED(a, "k_e_update<" #STORE ", " #MUS ", false>", (k_e_update<STORE, MUS>), gc, blk, 0, c->stream, \
                  c->m, c->o, cs, c->bc, Uin, Uout, sxy, dtp, dtr, bcv, a->d_dt, T, n, c->prm.t_max, P, Fb, \
                  nullptr, gs)
    if (muscl) {
        // MUSCL-EVR: the update from the fluxes of one edge pass (the dry
        // run's above; the replay's here: reconstruction with the step's own
        // dt and the recorded boundary values, boundary fluxes also into the
        // segment's slot Fb for the reverse sweep)
        if (!dry)
            KLAUNCH(a, k_e_edges_muscl, dim3(nblocks(c->ni + c->nb)), blk, 0, c->stream, c->m, c->o, cs, c->bc,
                    Uin, a->d_F, nullptr, nullptr, n, c->prm.t_max, sxy, nullptr, dtr, bcv, Fb);
        if (P)
            KLAUNCH_NAMED(a, "k_e_update<true, true, true>", (k_e_update<true, true, true>), gc, blk, 0,
                          c->stream, c->m, c->o, cs, c->bc, Uin, Uout, sxy, dtp, dtr, bcv, a->d_dt, T, n,
                          c->prm.t_max, P, Fb, a->d_F, gs);
        else
            KLAUNCH_NAMED(a, "k_e_update<false, true, true>", (k_e_update<false, true, true>), gc, blk, 0,
                          c->stream, c->m, c->o, cs, c->bc, Uin, Uout, sxy, dtp, dtr, bcv, a->d_dt, T, n,
                          c->prm.t_max, P, Fb, a->d_F, gs);
    } else if (P) {
        E_UPDATE(true, false);
    } else {
        E_UPDATE(false, false);
    }
#undef E_UPDATE
    if (dry && a->d_oframes)
        KLAUNCH(a, k_frame_sample, dim3(nblocks(3 * c->nc)), blk, 0, c->stream, Uout, c->nc, a->d_trig,
                n, a->d_oframes);
}

void launch_cost(sw2dg_ctx *c, AssimState *a, const double *U, int oi) {
    const double norm = (a->normalize && a->n_obs > 0) ? (double)a->n_obs : 1.0;
    if (a->n_gauges > 0)
        KLAUNCH(a, k_cost_gauges, dim3(1), dim3(64), 0, c->stream, U, a->d_gcell, a->d_gw,
                           a->d_gobs + (size_t)oi * a->n_gauges, a->n_gauges, norm, a->d_J + 1 + oi);
    if (a->frames) {
        const size_t o = (size_t)oi * c->nc;
        const int nb = nblocks(c->nc);
        KLAUNCH(a, k_cost_frame_part, dim3(nb), dim3(kBlock), 0, c->stream, U, (int)c->nc,
                           a->d_fh + o, a->d_fqx + o, a->d_fqy + o, a->d_wh, a->d_wq, a->d_cpart);
        KLAUNCH(a, k_cost_frame_sum, dim3(1), dim3(64), 0, c->stream, a->d_cpart, nb, norm,
                           a->d_J + 1 + oi);
    }
}

SeedArgs seed_args(sw2dg_ctx *c, AssimState *a, int oi, const double *U) {
    SeedArgs sd{};
    sd.oi = -1;
    if (oi < 0) return sd;
    sd.oi = oi;
    sd.U = U;
    sd.norm = (a->normalize && a->n_obs > 0) ? (double)a->n_obs : 1.0;
    if (a->n_gauges > 0) {
        sd.gfirst = a->d_gfirst;
        sd.gnext = a->d_gnext;
        sd.gw = a->d_gw;
        sd.gobs = a->d_gobs + (size_t)oi * a->n_gauges;
    }
    if (a->frames) {
        const size_t o = (size_t)oi * c->nc;
        sd.fh = a->d_fh + o;
        sd.fqx = a->d_fqx + o;
        sd.fqy = a->d_fqy + o;
        sd.wh = a->d_wh;
        sd.wq = a->d_wq;
    }
    return sd;
}

// segments, window and tiles (get_spatial_gradient_checkpointed, :785-953)
struct Plan {
    int64_t K = 0, first = 0;
    std::vector<char> tile_top;
    std::vector<int64_t> tile_start;
};

Plan make_plan(const AssimState *a, const std::vector<double> &T) {
    Plan p;
    const int64_t S = a->stride;
    p.K = (a->n_steps + S - 1) / S;
    p.first = 0;
    if (a->t_window > -1e299) {
        for (int64_t k = 0; k < p.K; ++k) {
            if (T[k * S] <= a->t_window + 1e-9)
                p.first = k;
            else
                break;
        }
    }
    // a frozen truncation window overrides the window's start (swSpE_checkpoint.cpp:878-881)
    if (a->forced_seg >= 0) p.first = std::min<int64_t>(a->forced_seg, std::max<int64_t>(p.K - 1, 0));
    p.tile_top.assign(p.K, 0);
    const int64_t M = p.K - p.first;
    int64_t n = std::max<int64_t>(1, std::min<int64_t>(a->n_tiles, M));
    for (int64_t t = 0; t < n; ++t) {
        const int64_t s_t = p.first + (t * M) / n;
        const int64_t e_t = p.first + ((t + 1) * M) / n - 1;
        p.tile_start.push_back(s_t);
        p.tile_top[e_t] = 1;
    }
    // after an adjoint explosion the cost starts where the gradient stopped
    // (replay_trajectory_cost, swSpE_checkpoint.cpp:1311-1400): tiles that
    // start before the kept segments are dropped, one tile otherwise
    if (a->trunc_seg >= 0 && a->trunc_seg + 1 < p.K) {
        const int64_t s0 = a->trunc_seg + 1;
        if (a->n_tiles > 1) {
            std::vector<int64_t> kept;
            for (int64_t st : p.tile_start)
                if (st >= s0) kept.push_back(st);
            p.tile_start = kept.empty() ? std::vector<int64_t>{s0} : kept;
        } else {
            p.tile_start = {s0};
        }
        p.first = std::max(p.first, s0);
    }
    return p;
}

int fetch_record(sw2dg_ctx *c, AssimState *a, std::vector<double> &T, std::vector<int> &trig,
                 std::vector<double> &dt) {
    T.resize(a->n_steps + 1);
    trig.resize(a->n_steps);
    dt.resize(a->n_steps);
    HIPCHK(c, hipMemcpy(T.data(), a->d_T, T.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (a->n_steps) {
        HIPCHK(c, hipMemcpy(trig.data(), a->d_trig, trig.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(dt.data(), a->d_dt, dt.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    a->h_dt = dt;
    return SW2DG_OK;
}

// replay segment k from its checkpoint into the segment buffer; cost of its
// observation steps added to J when with_cost
// replay segment k from its checkpoint (or `start`) into the segment
// buffers; store = keep what the reverse sweep reads (balanced states)
int replay_segment(sw2dg_ctx *c, AssimState *a, int64_t k, const std::vector<int> &trig,
                   bool with_cost, const double *start, bool store = false) {
    a->seg_blk = -1;  // the segment buffers, never a resident block
    const int64_t S = a->stride;
    const int64_t n0 = k * S, n1 = std::min(a->n_steps, n0 + S);
    HIPCHK(c, hipMemcpyAsync(seg_state(a, c, 0), start ? start : ckpt_state(a, c, k),
                             3 * c->nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    for (int64_t n = n0; n < n1; ++n) {
        const int j = (int)(n - n0);
        launch_e_step(c, a, seg_state(a, c, j), seg_state(a, c, j + 1), n, false, seg_flux(a, c, j),
                      store ? seg_pre(a, c, j) : nullptr,
                      (store && c->o.scheme == 1) ? seg_slopes(a, c, j) : nullptr);
        if (with_cost && trig[n] >= 0) launch_cost(c, a, seg_state(a, c, j + 1), trig[n]);
    }
    return SW2DG_OK;
}

}  // namespace

extern "C" {

int sw2dg_assim_set_phi_transform(sw2dg_ctx *c, int32_t mode, double lambda, const int64_t *free_cells,
                                  int64_t n_free, int32_t mask_wall) {
    if (!c || !c->assim || mode < 0 || mode > 3 || !(lambda > 0.0) || n_free < 0 ||
        (n_free > 0 && !free_cells))
        return SW2DG_ERR_ARG;
    for (int64_t k = 0; k < n_free; ++k)
        if (free_cells[k] < 0 || free_cells[k] >= c->nc) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    a->phi_mode = mode;
    a->phi_lambda = lambda;
    a->free_idx.assign(free_cells, free_cells + n_free);
    a->mask_wall = mask_wall != 0;
    a->norm_lo = 0.0;
    a->norm_scale = 1.0;
    return SW2DG_OK;
}

int sw2dg_phi_transform_eval(int32_t mode, double lambda, double lo, double hi, int64_t n,
                             const double *m, double *phi, double *dphi) {
    if (mode < 0 || mode > 3 || n < 0 || (n > 0 && !m)) return SW2DG_ERR_ARG;
    const double scale = (hi - lo < 1e-12) ? 1.0 : 1.0 / (hi - lo);
    const double l0 = (hi - lo < 1e-12) ? 0.0 : lo;
    for (int64_t i = 0; i < n; ++i) {
        if (phi) phi[i] = phi_correct(mode, lambda, l0, scale, m[i]);
        if (dphi) dphi[i] = dphi_dm(mode, lambda, scale, m[i]);
    }
    return SW2DG_OK;
}

int sw2dg_assim_set_explosion_guard(sw2dg_ctx *c, double ratio) {
    if (!c || !c->assim || !(ratio == ratio)) return SW2DG_ERR_ARG;
    c->assim->expl_ratio = ratio;
    return SW2DG_OK;
}

int sw2dg_assim_truncation(sw2dg_ctx *c, int64_t *first_kept_step, int64_t *first_kept_segment) {
    if (!c || !c->assim) return SW2DG_ERR_ARG;
    const AssimState *a = c->assim;
    if (first_kept_segment) *first_kept_segment = a->trunc_seg + 1;
    if (first_kept_step) *first_kept_step = (a->trunc_seg + 1) * a->stride;
    return SW2DG_OK;
}

int sw2dg_assim_set_first_segment(sw2dg_ctx *c, int64_t segment) {
    if (!c || !c->assim) return SW2DG_ERR_ARG;
    c->assim->forced_seg = segment >= 0 ? segment : -1;
    return SW2DG_OK;
}

int sw2dg_assim_set_adjoint_cut(sw2dg_ctx *c, double h_cut) {
    if (!c || !(h_cut >= 0.0)) return SW2DG_ERR_ARG;
    c->o.adj_cut = h_cut;
    return SW2DG_OK;
}

int sw2dg_assim_setup(sw2dg_ctx *c, const sw2dg_obs *obs, int32_t stride, int32_t n_tiles,
                      double t_window_start, int64_t max_steps) {
    if (!c || !obs || max_steps <= 0 || n_tiles < 1 || stride < 0) return SW2DG_ERR_ARG;
    if (!c->state_set) return fail(c, SW2DG_ERR_ARG, "sw2dg_set_state must be called first");
    if (c->o.wind || c->o.runoff || c->o.infil || c->o.mdiff)
        return fail(c, SW2DG_ERR_ARG,
                    "the 4D-Var layer has no wind, runoff, infiltration or momentum diffusion");
    if (obs->n_obs < 0 || obs->n_gauges < 0 || (obs->n_obs > 0 && !obs->t)) return SW2DG_ERR_ARG;
    if (obs->n_gauges > 0 && (!obs->gauge_cell || !obs->gauge_h)) return SW2DG_ERR_ARG;
    for (int g = 0; g < obs->n_gauges; ++g)
        if (obs->gauge_cell[g] < 0 || obs->gauge_cell[g] >= c->nc) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    assim_destroy(c->assim);
    c->assim = nullptr;
    auto *a = new AssimState();
    c->assim = a;
    a->stream = c->stream;
    const int64_t nc = c->nc, ni = c->ni, nb = c->nb;
    const int ns = std::max(c->n_series, 1);
    a->n_obs = obs->n_obs;
    a->n_gauges = obs->n_gauges;
    a->normalize = obs->normalize;
    a->obs_t.assign(obs->t, obs->t + obs->n_obs);
    a->max_steps = max_steps;
    a->stride = stride > 0 ? stride : (int)std::max<double>(1.0, std::ceil(std::sqrt((double)max_steps)));
    a->n_tiles = n_tiles;
    a->t_window = t_window_start;
    a->d_obs_t = a->alloc<double>(std::max(1, obs->n_obs));
    if (obs->n_obs)
        HIPCHK(c, hipMemcpy(a->d_obs_t, obs->t, obs->n_obs * sizeof(double), hipMemcpyHostToDevice));
    if (obs->n_gauges) {
        std::vector<int> gc(obs->n_gauges);
        for (int g = 0; g < obs->n_gauges; ++g) gc[g] = c->inv[obs->gauge_cell[g]];
        std::vector<double> gw(obs->n_gauges, 1.0);
        if (obs->gauge_w) gw.assign(obs->gauge_w, obs->gauge_w + obs->n_gauges);
        a->d_gcell = a->alloc<int>(obs->n_gauges);
        a->d_gw = a->alloc<double>(obs->n_gauges);
        a->d_gobs = a->alloc<double>((size_t)obs->n_gauges * std::max(1, obs->n_obs));
        a->d_gser = a->alloc<double>((size_t)obs->n_gauges * std::max(1, obs->n_obs));
        std::vector<int> gfirst(nc, -1), gnext(obs->n_gauges, -1), glast(nc, -1);
        for (int g = 0; g < obs->n_gauges; ++g) {
            const int cell = gc[g];
            if (glast[cell] < 0) gfirst[cell] = g;
            else gnext[glast[cell]] = g;
            glast[cell] = g;
        }
        a->d_gfirst = a->alloc<int>(nc);
        a->d_gnext = a->alloc<int>(obs->n_gauges);
        HIPCHK(c, hipMemcpy(a->d_gfirst, gfirst.data(), nc * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(a->d_gnext, gnext.data(), gnext.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(a->d_gcell, gc.data(), gc.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(a->d_gw, gw.data(), gw.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(a->d_gobs, obs->gauge_h,
                            (size_t)obs->n_gauges * obs->n_obs * sizeof(double), hipMemcpyHostToDevice));
    }
    if (obs->frame_h) {
        if (!obs->frame_qx || !obs->frame_qy) return fail(c, SW2DG_ERR_ARG, "frame observations incomplete");
        a->frames = true;
        const size_t tot = (size_t)nc * std::max(1, obs->n_obs);
        a->d_fh = a->alloc<double>(tot);
        a->d_fqx = a->alloc<double>(tot);
        a->d_fqy = a->alloc<double>(tot);
        std::vector<double> v(nc);
        const double *src[3] = {obs->frame_h, obs->frame_qx, obs->frame_qy};
        double *dst[3] = {a->d_fh, a->d_fqx, a->d_fqy};
        for (int f = 0; f < 3; ++f)
            for (int o = 0; o < obs->n_obs; ++o) {
                v = c->to_dev_cells(src[f] + (size_t)o * nc);
                HIPCHK(c, hipMemcpy(dst[f] + (size_t)o * nc, v.data(), nc * sizeof(double),
                                    hipMemcpyHostToDevice));
            }
        std::vector<double> one(nc, 1.0);
        a->d_wh = a->alloc<double>(nc);
        a->d_wq = a->alloc<double>(nc);
        std::vector<double> wh = obs->w_h ? c->to_dev_cells(obs->w_h) : one;
        std::vector<double> wq = obs->w_q ? c->to_dev_cells(obs->w_q) : one;
        HIPCHK(c, hipMemcpy(a->d_wh, wh.data(), nc * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(a->d_wq, wq.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    }
    a->d_T = a->alloc<double>(max_steps + 1);
    a->d_dt = a->alloc<double>(max_steps);
    a->d_bc = a->alloc<double>((size_t)max_steps * ns);
    a->d_trig = a->alloc<int>(max_steps);
    a->d_ck = a->alloc<EClock>(1);
    a->n_ckpt = (max_steps + a->stride - 1) / a->stride + 1;
    a->d_ckpt = a->alloc<double>((size_t)a->n_ckpt * 3 * nc);
    a->d_seg = a->alloc<double>((size_t)(a->stride + 1) * 3 * nc);
    a->d_Fseg = a->alloc<double>((size_t)a->stride * flux_len(c));
    a->d_Pseg = a->alloc<double>((size_t)a->stride * 6 * (size_t)nc);
    const bool muscl = c->o.scheme == 1;
    a->d_sxy = a->alloc<double>(muscl ? 2 * (size_t)nc : 1);
    a->d_F = a->alloc<double>(muscl ? 5 * (size_t)std::max<int64_t>(ni, 1) + 3 * (size_t)std::max<int64_t>(nb, 1) : 1);
    a->d_sw = a->alloc<double>(muscl ? 2 * (size_t)std::max<int64_t>(ni, 1) : 1);
    a->d_Sseg = a->alloc<double>(muscl ? (size_t)a->stride * 2 * (size_t)nc : 1);
    a->d_sd = a->alloc<double>(muscl ? 4 * (size_t)std::max<int64_t>(ni, 1) : 1);
    a->d_zn = a->alloc<double>(muscl ? 2 * (size_t)std::max<int64_t>(ni, 1) : 1);
    a->d_partial = a->alloc<double>(nblocks(nc));
    a->d_cpart = a->alloc<double>(nblocks(nc));
    a->d_J = a->alloc<double>(1 + (size_t)std::max(1, obs->n_obs));
    a->d_lam = a->alloc<double>(3 * nc);
    a->d_lams = a->alloc<double>(3 * nc);
    a->d_part = a->alloc<double>(3 * nc);
    a->d_side = a->alloc<double>(8 * (size_t)std::max<int64_t>(ni, 1));
    a->d_gphi = a->alloc<double>(nc);
    a->d_gK1 = a->alloc<double>(nc);
    a->d_gK2 = a->alloc<double>(nc);
    a->d_gbce = a->alloc<double>(std::max<int64_t>(nb, 1));
    a->d_gbc = a->alloc<double>((size_t)max_steps * ns);
    a->d_wm = a->alloc<double>(3 * nc);
    a->d_nrm = a->alloc<double>(nblocks(nc) + 1);
    a->d_phi = a->alloc<double>(nc);
    a->d_coe = a->alloc<double>(2 * (size_t)std::max<int64_t>(c->ni, 1));
    a->d_K1 = a->alloc<double>(nc);
    a->d_K2 = a->alloc<double>(nc);
    // boundary edges of each series in reference order (renumbered indices)
    std::vector<int> sp(ns + 1, 0), se;
    for (int k = 0; k < c->n_series; ++k) {
        for (int64_t e = 0; e < nb; ++e)
            if (c->bser_host[e] == k) se.push_back(c->binv[e]);
        sp[k + 1] = (int)se.size();
    }
    if (se.empty()) se.push_back(0);
    a->d_ser_ptr = a->alloc<int>(sp.size());
    a->d_ser_edges = a->alloc<int>(se.size());
    HIPCHK(c, hipMemcpy(a->d_ser_ptr, sp.data(), sp.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(a->d_ser_edges, se.data(), se.size() * sizeof(int), hipMemcpyHostToDevice));
    for (void *p : a->allocs)
        if (!p) return fail(c, SW2DG_ERR_HIP, "out of device memory for the adjoint buffers");
    return SW2DG_OK;
}

// the dry run; keep = it precedes a reverse sweep at the same controls: keep
// the resident trajectory when it fits
static int record_run(sw2dg_ctx *c, const double *phi, const double *K1, const double *K2,
                      int64_t *n_steps, bool keep) {
    if (n_steps) *n_steps = 0;
    if (!c || !c->assim) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    a->seg_blk = -1;
    a->res_kept.clear();
    HIPCHK(c, hipSetDevice(c->device));
    int st = upload_ctl(c, a, phi, K1, K2, /*new_bounds=*/true);
    if (st) return st;
    const int64_t nc = c->nc;
    // initial state -> checkpoint 0 and the dry-run buffers
    std::vector<double> U0(3 * nc);
    std::copy(c->h0.begin(), c->h0.end(), U0.begin());
    std::copy(c->qx0.begin(), c->qx0.end(), U0.begin() + nc);
    std::copy(c->qy0.begin(), c->qy0.end(), U0.begin() + 2 * nc);
    HIPCHK(c, hipMemcpy(ckpt_state(a, c, 0), U0.data(), 3 * nc * sizeof(double), hipMemcpyHostToDevice));
    EClock ek{};
    ek.oi = 0;
    ek.n_obs = a->n_obs;
    ek.n_series = c->n_series;
    ek.t_max = c->prm.t_max;
    ek.dt_max = c->prm.dt_max;
    ek.cfl = c->prm.cfl_max;
    HIPCHK(c, hipMemcpy(a->d_ck, &ek, sizeof(EClock), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(a->d_T, &c->t0, sizeof(double), hipMemcpyHostToDevice));
    if (a->n_gauges > 0) zero(c, a->d_gser, (size_t)a->n_gauges * std::max(1, a->n_obs));
    const int64_t S = a->stride;
    int64_t n = 0;
    int64_t k = 0;
    double tk = c->t0;
    while (n < a->max_steps && tk < c->prm.t_max) {
        // one segment: ping-pong in the segment buffer (or along its resident
        // block, keeping what the reverse sweep reads), then checkpoint its end
        const bool res = keep && resident_block(c, a, k) != nullptr;
        a->seg_blk = res ? k : -1;
        a->res_kept.push_back(res ? 1 : 0);
        HIPCHK(c, hipMemcpyAsync(seg_state(a, c, 0), ckpt_state(a, c, k), 3 * nc * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
        const int64_t n1 = std::min(a->max_steps, n + S);
        for (int64_t m2 = n; m2 < n1; ++m2) {
            const int j = (int)(m2 - n);
            launch_e_step(c, a, seg_state(a, c, j), seg_state(a, c, j + 1), m2, true,
                          res ? seg_flux(a, c, j) : nullptr, res ? seg_pre(a, c, j) : nullptr);
        }
        HIPCHK(c, hipMemcpyAsync(ckpt_state(a, c, k + 1), seg_state(a, c, (int)(n1 - n)),
                                 3 * nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        a->seg_blk = -1;
        HIPCHK(c, hipMemcpyAsync(&tk, a->d_T + n1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        n = n1;
        ++k;
    }
    // number of active steps: first n with T[n] >= t_max
    std::vector<double> T(n + 1);
    HIPCHK(c, hipMemcpy(T.data(), a->d_T, (n + 1) * sizeof(double), hipMemcpyDeviceToHost));
    int64_t cnt = 0;
    while (cnt < n && T[cnt] < c->prm.t_max) ++cnt;
    a->n_steps = cnt;
    a->recorded = true;
    a->seg_blk = -1;
    if (n_steps) *n_steps = cnt;  // < t_max reached when max_steps cut the window
    return SW2DG_OK;
}

int sw2dg_assim_record(sw2dg_ctx *c, const double *phi, const double *K1, const double *K2,
                       int64_t *n_steps) {
    return record_run(c, phi, K1, K2, n_steps, false);
}

int sw2dg_assim_set_resident(sw2dg_ctx *c, int32_t mode) {
    if (!c || !c->assim || mode < 0 || mode > 1) return SW2DG_ERR_ARG;
    c->assim->resident_mode = mode;
    c->assim->res_kept.clear();
    return SW2DG_OK;
}

int sw2dg_assim_resident(sw2dg_ctx *c, int32_t *active) {
    if (!c || !c->assim || !active) return SW2DG_ERR_ARG;
    int32_t n = 0;
    for (char v : c->assim->res_kept) n += v ? 1 : 0;
    *active = n;
    return SW2DG_OK;
}

int sw2dg_assim_get_record(sw2dg_ctx *c, double *t_seq, double *dt_seq, double *bc_seq,
                           int32_t *obs_trig) {
    if (!c || !c->assim || !c->assim->recorded) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = a->n_steps;
    const int ns = std::max(c->n_series, 1);
    if (t_seq) HIPCHK(c, hipMemcpy(t_seq, a->d_T, (n + 1) * sizeof(double), hipMemcpyDeviceToHost));
    if (dt_seq && n) HIPCHK(c, hipMemcpy(dt_seq, a->d_dt, n * sizeof(double), hipMemcpyDeviceToHost));
    if (bc_seq && n && c->n_series)
        HIPCHK(c, hipMemcpy(bc_seq, a->d_bc, (size_t)n * ns * sizeof(double), hipMemcpyDeviceToHost));
    if (obs_trig && n)
        HIPCHK(c, hipMemcpy(obs_trig, a->d_trig, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return SW2DG_OK;
}

int sw2dg_assim_set_bc_seq(sw2dg_ctx *c, const double *bc_seq) {
    if (!c || !c->assim || !c->assim->recorded || !bc_seq) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    const int ns = std::max(c->n_series, 1);
    if (a->n_steps && c->n_series)
        HIPCHK(c, hipMemcpy(a->d_bc, bc_seq, (size_t)a->n_steps * ns * sizeof(double),
                            hipMemcpyHostToDevice));
    return SW2DG_OK;
}

int sw2dg_assim_cost(sw2dg_ctx *c, const double *phi, const double *K1, const double *K2,
                     double *J) {
    if (!c || !c->assim || !c->assim->recorded || !J) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    int st = upload_ctl(c, a, phi, K1, K2);
    if (st) return st;
    std::vector<double> T, dt;
    std::vector<int> trig;
    if ((st = fetch_record(c, a, T, trig, dt))) return st;
    const Plan p = make_plan(a, T);
    zero(c, a->d_J, 1 + (size_t)std::max(1, a->n_obs));
    // tiles restart from their frozen checkpoints; segments inside a tile chain
    for (size_t ti = 0; ti < p.tile_start.size(); ++ti) {
        const int64_t s_t = p.tile_start[ti];
        const int64_t e_t = (ti + 1 < p.tile_start.size()) ? p.tile_start[ti + 1] - 1 : p.K - 1;
        for (int64_t k = s_t; k <= e_t; ++k) {
            const double *start = nullptr;
            if (k > s_t) {
                // continue from the end of the previous (full) segment; copy it out
                // first because the segment buffer is reused
                HIPCHK(c, hipMemcpyAsync(a->d_part, seg_state(a, c, a->stride), 3 * c->nc * sizeof(double),
                                         hipMemcpyDeviceToDevice, c->stream));
                start = a->d_part;
            }
            if ((st = replay_segment(c, a, k, trig, true, start))) return st;
        }
    }
    KLAUNCH(a, k_cost_total, dim3(1), dim3(64), 0, c->stream, a->d_J, a->n_obs);
    HIPCHK(c, hipMemcpyAsync(J, a->d_J, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SW2DG_OK;
}

int sw2dg_assim_gradient(sw2dg_ctx *c, const double *phi, const double *K1, const double *K2,
                         double *J, double *dJ_dphi, double *dJ_dK1, double *dJ_dK2,
                         double *dJ_dbc, int64_t *n_steps) {
    if (n_steps) *n_steps = 0;
    if (!c || !c->assim) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    int64_t nst = 0;
    int st = record_run(c, phi, K1, K2, &nst, true);
    if (st) return st;
    std::vector<double> T, dt;
    std::vector<int> trig;
    if ((st = fetch_record(c, a, T, trig, dt))) return st;
    a->trunc_seg = -1;  // reset from any previous gradient (swSpE_checkpoint.cpp:731)
    const Plan p = make_plan(a, T);
    // a window or a frozen segment that starts later is a truncation too: the
    // cost restarts there (swSpE_checkpoint.cpp:850-852, 878-881)
    if (p.first > 0) a->trunc_seg = p.first - 1;
    const int64_t nc = c->nc;
    const int ns = std::max(c->n_series, 1);
    zero(c, a->d_J, 1 + (size_t)std::max(1, a->n_obs));
    zero(c, a->d_lam, 3 * nc);
    zero(c, a->d_gphi, nc);
    zero(c, a->d_gK1, nc);
    zero(c, a->d_gK2, nc);
    zero(c, a->d_gbc, (size_t)std::max<int64_t>(a->n_steps, 1) * ns);
    const Controls cs = ctl(a);
    const dim3 blk(kBlock);
    const int ncb = nblocks(nc);
    const bool muscl = c->o.scheme == 1;
    const double *zn = muscl ? a->d_zn : nullptr;
    // fused gather (step n) + cell reverse (step n-1) [+ boundary-value sums of step n]
    auto gc = [&](bool cell, int gather, int64_t bc_step, const SeedArgs &sd, const double *P,
                  double dtv) {
        const int extra = (bc_step >= 0) ? c->n_series : 0;
        double *gbcn = (bc_step >= 0) ? a->d_gbc + (size_t)bc_step * ns : nullptr;
        if (cell)
            KLAUNCH(a, k_r_gc<true>, dim3(ncb + extra), blk, 0, c->stream, c->m, c->o, cs, a->d_part,
                    a->d_side, gather, sd, P, dtv, a->d_lams, a->d_lam, a->d_gphi, a->d_gK1, a->d_gK2,
                    ncb, a->d_ser_ptr, a->d_ser_edges, a->d_gbce, gbcn, zn);
        else
            KLAUNCH(a, k_r_gc<false>, dim3(ncb + extra), blk, 0, c->stream, c->m, c->o, cs, a->d_part,
                    a->d_side, gather, sd, P, dtv, a->d_lams, a->d_lam, a->d_gphi, a->d_gK1, a->d_gK2,
                    ncb, a->d_ser_ptr, a->d_ser_edges, a->d_gbce, gbcn, zn);
    };
    // edge reverse of step n (segment slot j); MUSCL-EVR: then the slopes
    // edge reverse of step n (with its boundary part in extra blocks)
    auto edges = [&](int64_t n, int j) {
        const int neb = nblocks(c->ni);
        const int nbb = c->m.n_bnd_cells > 0 ? (nblocks(c->m.n_bnd_cells) + 7) / 8 * 8 : 0;
        const BndRev br{c->bc, seg_flux(a, c, j), a->d_bc + (size_t)n * ns, a->d_part, a->d_gphi,
                        a->d_gbce, nbb};
        if (muscl) {
            KLAUNCH(a, k_r_edge<true>, dim3(neb + nbb), blk, 0, c->stream, c->m, c->o, cs,
                    seg_state(a, c, j), dt[n], a->d_lams, a->d_side, seg_slopes(a, c, j), a->d_sd, br);
            KLAUNCH(a, k_r_slopes, dim3(ncb), blk, 0, c->stream, c->m, seg_state(a, c, j), a->d_sd,
                    a->d_part, a->d_lams, a->d_zn);
        } else {
            KLAUNCH(a, k_r_edge<false>, dim3(neb + nbb), blk, 0, c->stream, c->m, c->o, cs,
                    seg_state(a, c, j), dt[n], a->d_lams, a->d_side, nullptr, nullptr, br);
        }
    };
    // explosion guard (swSpE_checkpoint.cpp:1035-1117): ||adjoint of h|| at
    // each segment start against the first non-zero one; a gradient snapshot
    // (watermark) at milestones (every 50th, first and last segment, segments
    // with observations) or, tiled, at each tile top; on explosion the
    // gradient returns to the watermark and the reverse sweep stops
    const bool guard = a->expl_ratio > 0.0 && a->forced_seg < 0;  // frozen window: no dynamic truncation
    const bool tiled = p.tile_start.size() > 1;
    const int nbk = nblocks(nc);
    double ref = -1.0;
    int64_t wm_seg = p.K - 1, wm_kept = p.K, wm_step = a->n_steps;
    auto snapshot = [&](int64_t from_step) -> int {
        HIPCHK(c, hipMemcpyAsync(a->d_wm, a->d_gphi, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(a->d_wm + nc, a->d_gK1, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(a->d_wm + 2 * nc, a->d_gK2, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        wm_step = from_step;
        return SW2DG_OK;
    };
    bool started = false;
    a->dh_peak = 0.0;
    for (int64_t k = p.K - 1; k >= p.first; --k) {
        if (guard && tiled && p.tile_top[k]) {  // clean state before a new tile
            if ((st = snapshot((k + 1) * a->stride))) return st;
            wm_kept = k + 1;
        }
        // forward recompute of segment k from its checkpoint (with its costs);
        // resident: the dry run kept it, only the costs are taken
        const int64_t n0 = k * a->stride, n1 = std::min(a->n_steps, n0 + a->stride);
        if (k < (int64_t)a->res_kept.size() && a->res_kept[k]) {
            a->seg_blk = k;
            for (int64_t n = n0; n < n1; ++n)
                if (trig[n] >= 0) launch_cost(c, a, seg_state(a, c, (int)(n - n0) + 1), trig[n]);
        } else if ((st = replay_segment(c, a, k, trig, true, nullptr, true))) {
            return st;
        }
        {
            const int64_t nt = n1 - 1;
            const SeedArgs sd = seed_args(c, a, trig[nt], seg_state(a, c, (int)(n1 - n0)));
            // the adjoint of the segment's end state was gathered when segment
            // k + 1 closed (into lam); a tile top drops it
            gc(true, (!started || p.tile_top[k]) ? -1 : 2, -1, sd, seg_pre(a, c, (int)(nt - n0)),
               dt[nt]);
            started = true;
        }
        for (int64_t n = n1 - 1; n >= n0; --n) {
            const int j = (int)(n - n0);
            edges(n, j);
            if (n > n0) {
                const SeedArgs sd = seed_args(c, a, trig[n - 1], seg_state(a, c, j));
                gc(true, 1, n, sd, seg_pre(a, c, j - 1), dt[n - 1]);
            }
        }
        // close the segment: adjoint of its start state (no seed: an
        // observation of that state belongs to segment k - 1), the porosity
        // terms of its first step and the boundary-value sums of step n0
        gc(false, 1, n0, seed_args(c, a, -1, nullptr), nullptr, 0.0);
        // ||adjoint of h|| at the segment start: its peak over the sweep is a
        // diagnostic of the reference's descent log (dh_peak, :1044-1049)
        KLAUNCH(a, k_sq_part, dim3(nbk), blk, 0, c->stream, a->d_lam, (int)nc, a->d_nrm);
        KLAUNCH(a, k_sq_sum, dim3(1), blk, 0, c->stream, a->d_nrm, nbk, a->d_nrm + nbk);
        double nrm2 = 0.0;
        HIPCHK(c, hipMemcpyAsync(&nrm2, a->d_nrm + nbk, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double nrm = std::sqrt(nrm2);
        if (nrm > a->dh_peak) a->dh_peak = nrm;
        if (!guard) continue;
        if (ref < 0.0 && nrm > 0.0) ref = nrm;
        if (!tiled) {
            bool has_obs = false;
            for (int64_t n = n0; n < n1; ++n) has_obs = has_obs || trig[n] >= 0;
            if (k % 50 == 0 || k == p.K - 1 || k == 0 || has_obs) {
                if ((st = snapshot(n0))) return st;
                wm_seg = k;
            }
        }
        if (ref > 0.0 && nrm > a->expl_ratio * ref) {
            HIPCHK(c, hipMemcpyAsync(a->d_gphi, a->d_wm, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(a->d_gK1, a->d_wm + nc, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(a->d_gK2, a->d_wm + 2 * nc, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            if (wm_step > 0)
                zero(c, a->d_gbc, (size_t)std::min(wm_step, a->n_steps) * ns);
            a->trunc_seg = (tiled ? wm_kept : wm_seg) - 1;
            // the cost covers what the gradient kept: observations of the
            // dropped segments leave the total
            const int64_t keep_from = (a->trunc_seg + 1) * a->stride;
            for (int64_t n = 0; n < std::min(keep_from, a->n_steps); ++n)
                if (trig[n] >= 0) zero(c, a->d_J + 1 + trig[n], 1);
            break;
        }
    }
    a->seg_blk = -1;
    KLAUNCH(a, k_cost_total, dim3(1), dim3(64), 0, c->stream, a->d_J, a->n_obs);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> v(nc);
    auto pull = [&](const double *src, double *dst) -> int {
        if (!dst) return SW2DG_OK;
        HIPCHK(c, hipMemcpy(v.data(), src, nc * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < nc; ++q) {
            const double g = v[q];
            // non-finite gradient entries are zeroed (swSpE_checkpoint.cpp:1197-1220)
            dst[c->perm[q]] = std::isfinite(g) ? g : 0.0;
        }
        return SW2DG_OK;
    };
    if ((st = pull(a->d_gphi, dJ_dphi))) return st;
    if (dJ_dphi && !a->m_last.empty()) {  // dJ/dm = dJ/dphi dphi/dm
        for (int64_t i = 0; i < nc; ++i) {
            const double m = a->m_last[i];
            dJ_dphi[i] *= dphi_dm(a->phi_mode, a->phi_lambda, a->norm_scale, m);
            // optional wall mask (mask_wall_gradient, swSpE_checkpoint.cpp:1206-1215)
            if (a->mask_wall &&
                !(phi_correct(a->phi_mode, a->phi_lambda, a->norm_lo, a->norm_scale, m) > c->o.phi_min))
                dJ_dphi[i] = 0.0;
        }
    } else if (dJ_dphi && a->mask_wall) {
        std::vector<double> ph(nc);
        HIPCHK(c, hipMemcpy(v.data(), a->d_phi, nc * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < nc; ++q) ph[c->perm[q]] = v[q];
        for (int64_t i = 0; i < nc; ++i)
            if (!(ph[i] > c->o.phi_min)) dJ_dphi[i] = 0.0;
    }
    if ((st = pull(a->d_gK1, dJ_dK1))) return st;
    if ((st = pull(a->d_gK2, dJ_dK2))) return st;
    if (dJ_dbc && a->n_steps && c->n_series)
        HIPCHK(c, hipMemcpy(dJ_dbc, a->d_gbc, (size_t)a->n_steps * ns * sizeof(double),
                            hipMemcpyDeviceToHost));
    if (J) HIPCHK(c, hipMemcpy(J, a->d_J, sizeof(double), hipMemcpyDeviceToHost));
    if (n_steps) *n_steps = a->n_steps;
    return SW2DG_OK;
}

int sw2dg_assim_profile(sw2dg_ctx *c, const double *phi, const double *K1, const double *K2,
                        int32_t k_max, double *avg_ms, int64_t *launches, const char **names,
                        int32_t *n_kernels) {
    if (!c || !c->assim || k_max < 0 || !n_kernels) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    a->prof = true;
    a->ev_used = 0;
    a->marks.clear();
    double J = 0.0;
    int64_t n = 0;
    const int st = sw2dg_assim_gradient(c, phi, K1, K2, &J, nullptr, nullptr, nullptr, nullptr, &n);
    a->prof = false;
    if (st) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<const char *> nm;
    std::vector<double> sum;
    std::vector<int64_t> cnt;
    for (const auto &mk : a->marks) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, a->ev[mk.second], a->ev[mk.second + 1]));
        size_t k = 0;
        while (k < nm.size() && std::strcmp(nm[k], mk.first) != 0) ++k;
        if (k == nm.size()) {
            nm.push_back(mk.first);
            sum.push_back(0.0);
            cnt.push_back(0);
        }
        sum[k] += ms;
        cnt[k] += 1;
    }
    a->marks.clear();
    *n_kernels = (int32_t)nm.size();
    for (size_t k = 0; k < nm.size() && (int32_t)k < k_max; ++k) {
        if (avg_ms) avg_ms[k] = sum[k] / (double)cnt[k];
        if (launches) launches[k] = cnt[k];
        if (names) names[k] = nm[k];
    }
    return SW2DG_OK;
}

int sw2dg_assim_gauge_series(sw2dg_ctx *c, double *out) {
    if (!c || !c->assim || !c->assim->recorded || !out) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    if (a->n_gauges == 0 || a->n_obs == 0) return SW2DG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(out, a->d_gser, (size_t)a->n_obs * a->n_gauges * sizeof(double),
                        hipMemcpyDeviceToHost));
    return SW2DG_OK;
}

int sw2dg_assim_adjoint_peak(sw2dg_ctx *c, double *dh_peak) {
    if (!c || !c->assim || !dh_peak) return SW2DG_ERR_ARG;
    *dh_peak = c->assim->dh_peak;
    return SW2DG_OK;
}

int sw2dg_assim_keep_obs_states(sw2dg_ctx *c, int32_t on) {
    if (!c || !c->assim) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    if (on && !a->d_oframes && a->n_obs > 0) {
        a->d_oframes = a->alloc<double>((size_t)a->n_obs * 3 * c->nc);
        if (!a->d_oframes) return fail(c, SW2DG_ERR_HIP, "keep_obs_states: out of device memory");
    }
    if (!on) a->d_oframes = nullptr;  // the buffer stays with the ctx until it is destroyed
    a->recorded = a->recorded && !on;  // the next dry run fills the frames
    return SW2DG_OK;
}

int sw2dg_assim_obs_states(sw2dg_ctx *c, double *h, double *qx, double *qy, int32_t *reached) {
    if (!c || !c->assim || !c->assim->recorded || !c->assim->d_oframes) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nc = c->nc;
    std::vector<int> trig(std::max<int64_t>(a->n_steps, 1));
    if (a->n_steps > 0)
        HIPCHK(c, hipMemcpy(trig.data(), a->d_trig, a->n_steps * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int32_t> hit(a->n_obs, 0);
    for (int64_t n = 0; n < a->n_steps; ++n)
        if (trig[n] >= 0 && trig[n] < a->n_obs) hit[trig[n]] = 1;
    std::vector<double> U(3 * nc);
    double *dst[3] = {h, qx, qy};
    for (int k = 0; k < a->n_obs; ++k) {
        if (reached) reached[k] = hit[k];
        if (!hit[k]) {
            for (int f = 0; f < 3; ++f)
                if (dst[f]) std::fill(dst[f] + (size_t)k * nc, dst[f] + (size_t)(k + 1) * nc, 0.0);
            continue;
        }
        HIPCHK(c, hipMemcpy(U.data(), a->d_oframes + (size_t)k * 3 * nc, 3 * nc * sizeof(double),
                            hipMemcpyDeviceToHost));
        for (int f = 0; f < 3; ++f)
            if (dst[f])
                for (int64_t q = 0; q < nc; ++q) dst[f][(size_t)k * nc + c->perm[q]] = U[f * nc + q];
    }
    return SW2DG_OK;
}

int sw2dg_assim_final_state(sw2dg_ctx *c, double *h, double *qx, double *qy) {
    if (!c || !c->assim || !c->assim->recorded) return SW2DG_ERR_ARG;
    AssimState *a = c->assim;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nc = c->nc;
    const int64_t K = (a->n_steps + a->stride - 1) / a->stride;
    std::vector<double> U(3 * nc);
    HIPCHK(c, hipMemcpy(U.data(), ckpt_state(a, c, K), 3 * nc * sizeof(double), hipMemcpyDeviceToHost));
    double *dst[3] = {h, qx, qy};
    for (int f = 0; f < 3; ++f)
        if (dst[f])
            for (int64_t q = 0; q < nc; ++q) dst[f][c->perm[q]] = U[f * nc + q];
    return SW2DG_OK;
}

}  // extern "C"
