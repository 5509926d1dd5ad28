// sw2dg_forward.hip -- MI355X (gfx950) forward step of the SW2D
// single-porosity model and the C-ABI declared in include/sw2dg.h.
//
// Step structure (reference: swSp::execute, swSp.cpp:522-628, and
// swSp::hypCompute, swSpHyp.cpp:137-211):
//   k_slopes  (MUSCL only)  per cell: z-gradient + in-place limiter replayed
//                           over the cell's internal edges in reference order
//   k_edges                 per internal edge: LR states, lambda_max, HLLC
//                           (only if one side is wet: dry-dry edges keep their
//                           previous contribution), contributions factor*flux
//                           for both sides, CFL swept terms
//   k_cfl                   per cell: swept = fixed-order sum, dt candidate,
//                           wave/block min -> per-block partial
//   k_dt                    one block: min of the partials, dt landing, the
//                           clock and the boundary values
//   k_update                per cell: fixed-order gather of boundary fluxes and
//                           edge contributions (no atomics: deterministic and
//                           in the reference's summation order), round-off
//                           clean-up, negative-depth clamp, friction, singular
//                           head loss, velocity limiter
// Cells are renumbered along a Hilbert curve, internal edges sorted by their
// lower renumbered cell, and every launch maps workgroups to XCDs in
// contiguous ranges so neighbouring blocks share an L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"
#include "sp_physics.h"
#include "sw2dg_internal.h"

using sp::kG;
using sp::kSmall;
using sp::smax;
using sp::smin;

using namespace sw2dg_detail;

namespace {

constexpr int kStepsPerGraph = 32;  // even: the clock double buffer returns to slot 0
constexpr int kMaxParabSubsteps = 1 << 20;  // explicit diffusion substeps of one step, at most

__device__ __forceinline__ bool step_active(const Clock *ck, int p) {
    return ck->ACT[p] != 0 && ck->err == 0;
}

__device__ __forceinline__ void raise_err(Clock *ck, int code) { atomicCAS(&ck->err, 0, code); }

// inactive step: block 0 / thread 0 carries the clock over to the other slot
__device__ __forceinline__ void carry_clock(Clock *ck, int p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ck->T[1 - p] = ck->T[p];
        ck->DT[1 - p] = ck->DT[p];
        ck->STEP[1 - p] = ck->STEP[p];
        ck->ACT[1 - p] = 0;
    }
}

// ---------------------------------------------------------------------------
// k_slopes: MUSCL-EVR free-surface gradient and limiter
// (swSp::computeMusclSlopes, swSpHypMusclEvr.cpp:11-115).  The reference
// limits in place edge by edge; a cell's slope is only modified by its own
// incident edges using its own current slope, so replaying the cell's edges
// in ascending reference order reproduces it exactly.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_slopes(MeshDev m, Clock *ck, int p) {
    if (!step_active(ck, p)) return;
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= m.n_cell) return;
    const double zi = m.h[i] + m.zb[i];
    const double xi = m.cx[i], yi = m.cy[i];
    const int b = m.int_ptr[i], e_end = m.int_ptr[i + 1];
    const int nk = e_end - b;
    // the first KC neighbours' free surface and incidence stay in registers
    // for the limiter pass (triangles and quads: all of them)
    constexpr int KC = 4;
    double zn[KC];
    int in[KC];
    double sx = 0.0, sy = 0.0, wsum = 0.0;
    auto grad_term = [&](int j, double &zj_out) {  // j: the neighbour (int_nbr)
        const double zj = m.h[j] + m.zb[j];
        zj_out = zj;
        if (nk > 1) {
            const double dx = m.cx[j] - xi;
            const double dy = m.cy[j] - yi;
            const double dist = sqrt(dx * dx + dy * dy);
            const double w = 1.0 / dist;
            sx += (zj - zi) * dx / dist / dist * w;
            sy += (zj - zi) * dy / dist / dist * w;
            wsum += w;
        }
    };
#pragma unroll
    for (int t = 0; t < KC; ++t) {
        if (t < nk) {
            in[t] = m.int_inc[b + t];
            grad_term(m.int_nbr[b + t], zn[t]);
        }
    }
    for (int k = b + KC; k < e_end; ++k) {
        double zj;
        grad_term(m.int_nbr[k], zj);
    }
    if (nk > 1) {
        sx /= wsum;
        sy /= wsum;
    }
    // limiter replayed over the incidences in order; (zL, zR) of the edge is
    // (zi, zj) or (zj, zi) by the side of this cell
    auto limit = [&](int inc, double zj) {
        const int e = inc >> 1;
        const double zL = (inc & 1) ? zj : zi;
        const double zR = (inc & 1) ? zi : zj;
        const double fmx = smax(zL, zR);
        const double fmn = smin(zL, zR);
        const double fr = zi + (m.exm[e] - xi) * sx + (m.eym[e] - yi) * sy;
        if (fr < fmn) {
            const double sf = (fmn - zi) / (fr - zi);
            sx *= sf;
            sy *= sf;
        }
        if (fr > fmx) {
            const double sf = (fmx - zi) / (fr - zi);
            sx *= sf;
            sy *= sf;
        }
    };
#pragma unroll
    for (int t = 0; t < KC; ++t)
        if (t < nk) limit(in[t], zn[t]);
    for (int k = b + KC; k < e_end; ++k) {
        const int j = m.int_nbr[k];
        limit(m.int_inc[k], m.h[j] + m.zb[j]);
    }
    m.dzdx[i] = sx;
    m.dzdy[i] = sy;
}

// Riemann state on one side of an internal edge.
struct Side {
    double z, q, r, lm, lp;
};

// Godunov: swSpHypGodunov.cpp:40-81
__device__ __forceinline__ Side side_godunov(double h, double qx, double qy, double zb, double b,
                                             double nx, double ny) {
    Side s;
    s.z = h + zb;
    s.q = qx * nx + qy * ny;
    s.r = -qx * ny + qy * nx;
    const double u = s.q / smax(h, kSmall);
    const double c = sqrt(kG * h);
    s.lm = sp::lambda_minus(b, c, u);
    s.lp = sp::lambda_plus(b, c, u);
    return s;
}

// MUSCL-EVR: swSpHypMusclEvr.cpp:250-352.  right = 0 for the left cell.
__device__ __forceinline__ Side side_muscl(bool right, double h, double qx, double qy, double zb,
                                           double b, double nx, double ny, double xm, double ym,
                                           double xc, double yc, double sx, double sy, double dt) {
    const double ql = qx * nx + qy * ny;
    const double rl = qy * nx - qx * ny;
    const double zc = h + zb;
    const double cl = sqrt(kG * smax(kSmall, h));
    const double ul = ql / smax(kSmall, h);
    const double vl = rl / smax(kSmall, h);
    double l1, l2;
    if (!right) {
        l1 = smax(0.0, sp::lambda_minus(b, cl, ul));
        l2 = smax(0.0, sp::lambda_plus(b, cl, ul));
    } else {
        l1 = smin(0.0, sp::lambda_minus(b, cl, ul));
        l2 = smin(0.0, sp::lambda_plus(b, cl, ul));
    }
    const double dx1 = xm - xc - 0.5 * l1 * dt * nx;
    const double dy1 = ym - yc - 0.5 * l1 * dt * ny;
    const double z1 = zc + dx1 * sx + dy1 * sy;
    const double h1 = smax(0.0, z1 - zb);
    const double dx2 = xm - xc - 0.5 * l2 * dt * nx;
    const double dy2 = ym - yc - 0.5 * l2 * dt * ny;
    const double z2 = zc + dx2 * sx + dy2 * sy;
    const double h2 = smax(0.0, z2 - zb);
    const double hr = 0.5 * (h1 + h2);
    Side s;
    s.z = 0.5 * (h1 + h2) + zb;
    s.q = 0.5 * (h1 + h2) * ul;
    s.r = 0.5 * (h1 + h2) * vl;
    const double u = s.q / smax(hr, kSmall);
    const double c = sqrt(kG * hr);
    s.lm = sp::lambda_minus(b, c, u);
    s.lp = sp::lambda_plus(b, c, u);
    return s;
}

// ---------------------------------------------------------------------------
// k_edges: one thread per internal edge (swSpHyp.cpp:318-400 for the flux,
// swSpHyp.cpp:454-477 for the CFL swept terms, swSp.cpp:405-441 for the
// edge factors, applied here so the cell gather only multiplies by dt).
// ---------------------------------------------------------------------------
template <int SCHEME>
__global__ __launch_bounds__(kBlock) void k_edges(MeshDev m, Opts o, Clock *ck, int p) {
    if (!step_active(ck, p)) return;
    const int e = xcd_block() * kBlock + (int)threadIdx.x;
    if (e >= m.n_int) return;
    const int L = m.eL[e], R = m.eR[e];
    const double nx = m.enx[e], ny = m.eny[e], len = m.elen[e];
    const double hL = m.h[L], hR = m.h[R];
    const double zbL = m.zb[L], zbR = m.zb[R];
    const double bL = m.beta[L], bR = m.beta[R];
    Side sL, sR;
    if (SCHEME == 0) {
        sL = side_godunov(hL, m.qx[L], m.qy[L], zbL, bL, nx, ny);
        sR = side_godunov(hR, m.qx[R], m.qy[R], zbR, bR, nx, ny);
    } else {
        const double dt = ck->DT[p];
        const double xm = m.exm[e], ym = m.eym[e];
        sL = side_muscl(false, hL, m.qx[L], m.qy[L], zbL, bL, nx, ny, xm, ym, m.cx[L], m.cy[L],
                        m.dzdx[L], m.dzdy[L], dt);
        sR = side_muscl(true, hR, m.qx[R], m.qy[R], zbR, bR, nx, ny, xm, ym, m.cx[R], m.cy[R],
                        m.dzdx[R], m.dzdy[R], dt);
    }
    double lmax = fabs(sL.lm);
    double v = fabs(sL.lp);
    if (lmax < v) lmax = v;
    v = fabs(sR.lm);
    if (lmax < v) lmax = v;
    v = fabs(sR.lp);
    if (lmax < v) lmax = v;
    // native CFL quirk: the lambda picked up on the left carries to the right.
    // Swept terms and contributions go to separate dense arrays with 16-byte
    // stores: the CFL gather then reads 16 bytes per edge, not the record.
    const bool wL = hL > o.h_min;
    const bool wR = hR > o.h_min;
    reinterpret_cast<double2 *>(m.swp)[e] =
        make_double2((wL ? lmax : 0.) * len, ((wL || wR) ? lmax : 0.) * len);
    if ((sL.z - zbL > o.h_min) || (sR.z - zbR > o.h_min)) {
        const double phiL = m.phi[L], phiR = m.phi[R];
        double f[5];
        const int st = sp::hllc(o.topo == 1, o.h_min, o.phi_min, bL, bR, phiL, phiR, zbL, zbR,
                                sL.z, sL.q, sL.r, sR.z, sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, f);
        if (st != sp::ST_OK) {
            raise_err(ck, st);
            return;
        }
        if (isnan(f[0])) {
            raise_err(ck, sp::ST_NAN);
            return;
        }
        const double cl = (phiL > o.phi_min) ? len / m.area[L] / phiL : 0.0;
        const double cr = (phiR > o.phi_min) ? len / m.area[R] / phiR : 0.0;
        const double a0 = (-cl) * f[0];
        const double a1 = ((-cl) * nx) * f[1] + (cl * nx) * f[3] + (cl * ny) * f[2];
        const double a2 = ((-cl) * ny) * f[1] + (cl * ny) * f[3] + ((-cl) * nx) * f[2];
        const double b0 = cr * f[0];
        const double b1 = (cr * nx) * f[1] + (cr * nx) * f[4] + ((-cr) * ny) * f[2];
        const double b2 = (cr * ny) * f[1] + (cr * ny) * f[4] + (cr * nx) * f[2];
        double2 *rec = reinterpret_cast<double2 *>(m.flx + 6 * (size_t)e);
        rec[0] = make_double2(a0, a1);
        rec[1] = make_double2(a2, b0);
        rec[2] = make_double2(b1, b2);
    }
}

// lambda_max of a boundary edge from its cell state (swSpHypGodunov.cpp:18-36)
__device__ __forceinline__ double bnd_lmax(double h, double qx, double qy, double b, double nx,
                                           double ny, double *lp_out, double *qr, double *rr) {
    const double q = qx * nx + qy * ny;
    const double r = -qx * ny + qy * nx;
    const double u = q / smax(h, kSmall);
    const double c = sqrt(kG * h);
    const double lm = sp::lambda_minus(b, c, u);
    const double lp = sp::lambda_plus(b, c, u);
    if (lp_out) *lp_out = lp;
    if (qr) *qr = q;
    if (rr) *rr = r;
    return smax(fabs(lp), fabs(lm));
}

// swBoundaryConditionTimeSeries::computeValBC (swBoundaryConditionTimeSeries.cpp:61-96)
__device__ void bc_advance(const BcDev &bc, int k, double time) {
    const double *v = bc.v + (size_t)k * bc.n_times;
    while ((time > bc.tsup[k]) && (bc.atend[k] == 0)) {
        bc.vinf[k] = bc.vsup[k];
        bc.tinf[k] = bc.tsup[k];
        bc.tsup[k] = bc.t[bc.rank[k]];
        bc.vsup[k] = v[bc.rank[k]];
        ++bc.rank[k];
        if (bc.rank[k] == bc.n_times) bc.atend[k] = 1;
    }
    if (time > bc.tsup[k]) {
        bc.val[k] = bc.vsup[k];
    } else if (time < bc.tinf[k]) {
        bc.val[k] = bc.vinf[k];
    } else {
        const double delta_t = (time - bc.tinf[k]) / (bc.tsup[k] - bc.tinf[k]);
        bc.val[k] = bc.vinf[k] + (bc.vsup[k] - bc.vinf[k]) * delta_t;
    }
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

// ---------------------------------------------------------------------------
// k_cfl: per-cell swept sum in reference order (boundary edges first, then
// internal edges, swSpHyp.cpp:440-477), dt candidate (:481-493), wavefront and
// block min -> one partial per block (min is exact, so the reduction order
// does not matter).  k_dt, one block, then reduces the partials and finalises
// the step's dt (map landing and t_max clamp, :495-508), the boundary values
// at the step's start time, the clock and the storage flags
// (swParent::checkStore, swParent.cpp:46-88).  The kernel boundary orders the
// two: no grid-wide fence or atomic (a device-scope release writes back the
// L2 of an XCD, which per block costs more than the whole reduction).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cfl(MeshDev m, Opts o, Clock *ck, int p) {
    if (!step_active(ck, p)) return;
    __shared__ double red[kBlock / 64];
    const int lb = xcd_block();
    const int i = lb * kBlock + (int)threadIdx.x;
    double cand = ck->dt_max;
    if (i < m.n_cell) {
        const double h = m.h[i];
        double swept = 0.;
        const int bb = m.bnd_ptr[i], be = m.bnd_ptr[i + 1];
        if (h > o.h_min) {
            for (int k = bb; k < be; ++k) {
                const int j = m.bnd_inc[k];
                const double lm =
                    bnd_lmax(h, m.qx[i], m.qy[i], m.beta[i], m.bnx[j], m.bny[j], nullptr, nullptr, nullptr);
                swept += lm * m.blen[j];
            }
        }
        const int ib = m.int_ptr[i], ie = m.int_ptr[i + 1];
        for (int kb = ib; kb < ie; kb += 4) {  // four incidences' loads in flight, summed in order
            int in[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) in[j] = (kb + j < ie) ? m.int_inc[kb + j] : 0;
            double w[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) w[j] = (kb + j < ie) ? m.swp[in[j]] : 0.0;  // (edge << 1) | side
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (kb + j < ie) swept += w[j];
        }
        const double x = ck->cfl * (m.area[i] / smax(swept, kSmall));
        if (x < cand) cand = x;
    }
    cand = wave_min(cand);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = cand;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < kBlock / 64; ++w) v = (red[w] < v) ? red[w] : v;
        m.partial[lb] = v;
    }
}

// Record bracket of swParent::interpolateTimeSeries (swParent.cpp:331-387)
// at time tw on non-decreasing record times: -1 at or before the first record
// less small_real (value of record 0), -2 at or after the last record (value
// of the last), else the first k >= 1 with tw <= t[k] (its sweep stops there;
// the value interpolates records k-1 and k).  A one-record table gives -1 in
// the sliver (t[0] - small_real, t[0]), where the reference reads past it.
__device__ int record_bracket(const double *tv, int n, double tw) {
    if (tw <= tv[0] - kSmall) return -1;
    if (tw >= tv[n - 1]) return -2;
    if (n < 2) return -1;
    int lo = 1, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (tw <= tv[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// momentum diffusion, velocity_independent (swSpParab.cpp:10-96): edge
// coefficient = harmonic mean of nu h phi of its cells, explicit substeps of
// dt / n with n = int(dt / min(t_max - t_0, dt, min over edge sides of
// area phi / (2 D))), on qx and qy.  The time-loop graph holds one substep;
// a step that needs n > 1 leaves them pending (k_parab_n): the rest of the
// graph's steps idle, and the host runs the n substeps and the end of the
// step before resuming (finish_parab).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double parab_coef(const MeshDev &m, int L, int R) {
    const double dL = sp::kMomDiff * m.h[L] * m.phi[L];
    const double dR = sp::kMomDiff * m.h[R] * m.phi[R];
    return (smax(dL, dR) > kSmall) ? 2.0 * dL * dR / (dL + dR) : 0.0;
}

// edge factor length / max(|d . n|, small_real) (swSp.cpp:1236-1246), once per mesh
__global__ __launch_bounds__(kBlock) void k_parab_fac(MeshDev m, double *fac) {
    const int e = (int)blockIdx.x * kBlock + (int)threadIdx.x;
    if (e >= m.n_int) return;
    const int L = m.eL[e], R = m.eR[e];
    const double dx = m.cx[R] - m.cx[L];
    const double dy = m.cy[R] - m.cy[L];
    const double d = fabs(dx * m.enx[e] + dy * m.eny[e]);
    fac[e] = m.elen[e] / smax(d, kSmall);
}

// per-block minimum of the substep bounds of the internal edges
__global__ __launch_bounds__(kBlock) void k_parab_dt(MeshDev m, Clock *ck, int p) {
    if (!step_active(ck, p)) return;
    __shared__ double red[kBlock / 64];
    const int e = xcd_block() * kBlock + (int)threadIdx.x;
    double v = __builtin_inf();
    if (e < m.n_int) {
        const int L = m.eL[e], R = m.eR[e];
        const double D = parab_coef(m, L, R);
        const double a = 0.5 * m.area[L] * m.phi[L] / D;
        const double b = 0.5 * m.area[R] * m.phi[R] / D;
        v = (a < v) ? a : v;  // a NaN bound (0 / 0) is never taken, as with std::min
        v = (b < v) ? b : v;
    }
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    v = red[0];
    for (int w = 1; w < kBlock / 64; ++w) v = (red[w] < v) ? red[w] : v;
    m.ppart[blockIdx.x] = v;
}

__global__ __launch_bounds__(kDtBlock) void k_parab_n(MeshDev m, Opts o, Clock *ck, int p, int n_part) {
    if (!step_active(ck, p)) return;
    __shared__ double red[kDtBlock / 64];
    double v = o.parab_span;
    for (int k = threadIdx.x; k < n_part; k += kDtBlock) {
        const double x = m.ppart[k];
        v = (x < v) ? x : v;
    }
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double dtp = red[0];
    for (int w = 1; w < kDtBlock / 64; ++w) dtp = (red[w] < dtp) ? red[w] : dtp;
    const double dt = ck->DT[1 - p];
    dtp = (dt < dtp) ? dt : dtp;
    const double ratio = dt / dtp;
    if (!(ratio < (double)kMaxParabSubsteps + 1.0)) {
        raise_err(ck, sp::ST_PARAB_SUBSTEPS);
        return;
    }
    const int n = (int)ratio;
    ck->parab_dt = dt / n;
    ck->parab_n = n;
    if (n > 1) {  // pending: the following steps of this launch idle
        ck->parab_pending = 1;
        ck->act_saved = ck->ACT[1 - p];
        ck->ACT[1 - p] = 0;
    }
}

// one explicit substep of qx and qy, (sx, sy) -> (dx, dy); each internal
// edge flux is formed by both of its cells from the same values, in the
// reference's edge order
__global__ __launch_bounds__(kBlock) void k_parab_sub(MeshDev m, Clock *ck, int p,
                                                      const double *__restrict__ sx,
                                                      const double *__restrict__ sy,
                                                      double *__restrict__ dx, double *__restrict__ dy) {
    if (!step_active(ck, p) || ck->parab_pending) return;
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= m.n_cell) return;
    const double dtp = ck->parab_dt;
    const double den = smax(kSmall, m.phi[i]);
    const double area = m.area[i];
    double vx = sx[i], vy = sy[i];
    const int ib = m.int_ptr[i], ie = m.int_ptr[i + 1];
    for (int k = ib; k < ie; ++k) {
        const int e = m.int_inc[k] >> 1;
        const int L = m.eL[e], R = m.eR[e];
        const double w = parab_coef(m, L, R) * m.pfac[e];
        const double hL = m.h[L], hR = m.h[R];
        const double fx = w * (sx[L] / hL - sx[R] / hR);
        const double fy = w * (sy[L] / hL - sy[R] / hR);
        if (i == L) {
            vx = vx - dtp * fx / area / den;
            vy = vy - dtp * fy / area / den;
        } else {
            vx = vx + dtp * fx / area / den;
            vy = vy + dtp * fy / area / den;
        }
    }
    dx[i] = vx;
    dy[i] = vy;
}

// Value of one station's records at time tw given its bracket k
// (record_bracket): the interpolation of swParent.cpp:383-384.
__device__ __forceinline__ double record_value(const double *tv, const double *v, int n, int k,
                                              double tw) {
    if (k < 1) return v[(k == -2) ? n - 1 : 0];
    const double t1 = tv[k - 1], t2 = tv[k];
    return v[k - 1] + (v[k] - v[k - 1]) * (tw - t1) / (t2 - t1);
}

__global__ __launch_bounds__(kDtBlock) void k_dt(MeshDev m, BcDev bc, Clock *ck, int p, int n_part) {
    if (!step_active(ck, p)) return;
    __shared__ double red[kDtBlock / 64];
    const double dt_max = ck->dt_max;
    double v = dt_max;
    for (int k0 = threadIdx.x; k0 < n_part; k0 += 4 * kDtBlock) {  // four loads in flight per thread
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = (k0 + j * kDtBlock < n_part) ? m.partial[k0 + j * kDtBlock] : dt_max;
#pragma unroll
        for (int j = 0; j < 4; ++j) v = (x[j] < v) ? x[j] : v;
    }
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double dt = red[0];
    for (int w = 1; w < kDtBlock / 64; ++w) dt = (red[w] < dt) ? red[w] : dt;
    const double t = ck->T[p];
    if (ck->dt_map > kSmall) {
        const int now = (int)(t / ck->dt_map);
        const int later = (int)((t + dt) / ck->dt_map);
        if (later > now) dt = later * ck->dt_map - t;
    }
    if ((t + dt) > ck->t_max) dt = ck->t_max - t;
    for (int k = 0; k < ck->n_series; ++k) bc_advance(bc, k, t);
    if (m.wind_nt > 0 || m.rain_nt > 0) {
        const double tw = t + 0.5 * dt;
        ck->t_half = tw;
        if (m.wind_nt > 0) ck->wind_k = record_bracket(m.wind_t, m.wind_nt, tw);
        if (m.rain_nt > 0) ck->rain_k = record_bracket(m.rain_t, m.rain_nt, tw);
    }
    const double tn = t + dt;
    int sf = 0;
    if (ck->dt_map <= -kSmall) {
        sf |= 1;
    } else if (ck->dt_map > kSmall) {
        if ((tn / ck->dt_map) == (double)(int)(tn / ck->dt_map)) sf |= 1;
    }
    if (ck->dt_ts <= -kSmall) {
        sf |= 2;
    } else if (ck->dt_ts > kSmall) {
        const int a = (int)((tn - dt) / ck->dt_ts);
        const int b2 = (int)(tn / ck->dt_ts);
        if (b2 > a) sf |= 2;
    }
    const long long step = ck->STEP[p] + 1;
    ck->store_flags = sf;
    ck->dt_last = dt;
    ck->T[1 - p] = tn;
    ck->DT[1 - p] = dt;
    ck->STEP[1 - p] = step;
    const bool pause = (ck->pause_on_store != 0) && (sf != 0);
    if (pause) ck->paused = 1;
    ck->ACT[1 - p] = (tn < ck->t_max) && (tn < ck->t_stop) && (step < ck->step_stop) && !pause;
}

// ---------------------------------------------------------------------------
// k_update: balance (swSpHyp.cpp:214-273) + clamp (:183-209) + friction
// (swSpSourceFriction.cpp:39-171) + singular head loss (:174-218) +
// correctVelocity (swSp.cpp:447-479), all local to one cell.
// ---------------------------------------------------------------------------
// part 0: the whole update; with momentum diffusion the parabolic step sits
// between the clamp and friction (swSp.cpp:535-545): part 1 stops after the
// clamp, part 2 starts from the diffused discharges.  EXTRA = false compiles
// the step without the runoff, infiltration, wind and diffusion options
// (part 0 only), which frees registers for more waves per SIMD.
template <bool EXTRA>
__global__ __launch_bounds__(kBlock) void k_update(MeshDev m, Opts o, BcDev bc, Clock *ck, int p,
                                                   int part_arg) {
    const int part = EXTRA ? part_arg : 0;
    if (!step_active(ck, p)) {
        carry_clock(ck, p);
        return;
    }
    if (part == 2 && ck->parab_pending) return;  // finished by the host
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= m.n_cell) return;
    const double dt = ck->DT[1 - p];
    const double h0 = m.h[i];
    const double qx0 = (part == 2) ? m.pqx[i] : m.qx[i];
    const double qy0 = (part == 2) ? m.pqy[i] : m.qy[i];
    double h = h0, qx = qx0, qy = qy0;
    if (part != 2) {
    const int bb = m.bnd_ptr[i], be = m.bnd_ptr[i + 1];
    if (be > bb) {
        const double zb = m.zb[i], phi = m.phi[i], b = m.beta[i];
        for (int k = bb; k < be; ++k) {
            const int j = m.bnd_inc[k];
            const double nx = m.bnx[j], ny = m.bny[j];
            double lp, qr, rr;
            bnd_lmax(h0, qx0, qy0, b, nx, ny, &lp, &qr, &rr);
            const int s = m.bser[j];
            double f[3];
            const int st = sp::bnd_flux(o.h_min, b, phi, zb, h0 + zb, qr, rr, lp, bc.type[s], bc.val[s], f);
            if (st != sp::ST_OK) {
                raise_err(ck, st);
                return;
            }
            m.bF1[j] = f[0];
            const double c = (phi > o.phi_min) ? m.blen[j] / m.area[i] / phi : 0.;
            h += c * f[0] * dt;
            qx += ((c * nx) * f[1] + (c * nx) * 0.0 + ((-c) * ny) * f[2]) * dt;
            qy += ((c * ny) * f[1] + (c * ny) * 0.0 + (c * nx) * f[2]) * dt;
        }
    }
    const int ib = m.int_ptr[i], ie = m.int_ptr[i + 1];
    for (int k = ib; k < ie; ++k) {
        const double *r = m.flx + 3 * (size_t)m.int_inc[k];  // 6 * edge + 3 * side
        h += r[0] * dt;
        qx += r[1] * dt;
        qy += r[2] * dt;
    }
    if (fabs(h) < kSmall) h = 0.;
    if (fabs(qx) < kSmall) qx = 0.;
    if (fabs(qy) < kSmall) qy = 0.;
    if (h < 0.) {
        if (o.neg_det) {
            raise_err(ck, sp::ST_NEG_DEPTH);
            return;
        }
        h = 0.;
        qx = 0.;
        qy = 0.;
    }
    }
    if (part == 1) {
        m.h[i] = h;
        m.qx[i] = qx;
        m.qy[i] = qy;
        return;
    }
    if (o.friction != sp::FRIC_NONE) {
        if (h > smax(o.h_min, kSmall)) {
            const double u = qx / h;
            const double v = qy / h;
            const double ca = m.fcos[i], sa = m.fsin[i];
            double ul = u * ca + v * sa;
            double vl = -u * sa + v * ca;
            const double un = sqrt(u * u + v * v);
            const double hp = (o.friction != sp::FRIC_CHEZY && h > kSmall) ? sp::friction_hpow(h) : 0.0;
            const double c1 = m.fc1[i], c2 = m.fc2[i];
            const double f1 = sp::friction_factor(o.friction, c1, h, un, dt, hp);
            const double f2 = (c2 == c1) ? f1 : sp::friction_factor(o.friction, c2, h, un, dt, hp);
            ul *= f1;
            vl *= f2;
            qx = (ul * ca - vl * sa) * h;
            qy = (ul * sa + vl * ca) * h;
        } else {
            qx = 0.;
            qy = 0.;
        }
    }
    if (o.head_loss) {
        const double ca = m.hcos[i], sa = m.hsin[i];
        double ax = qx * ca + qy * sa;
        double ay = -qx * sa + qy * ca;
        if (h > smax(o.h_min, kSmall)) {
            double aax = 0, aay = 0;
            if (h > m.hsing[i]) {
                aax = sqrt(ax * ax + ay * ay) * m.hk1[i] / h;
                aay = sqrt(ax * ax + ay * ay) * m.hk2[i] / h;
            }
            ax = ax * sw_exp(-aax * dt);
            ay = ay * sw_exp(-aay * dt);
        } else {
            ax = 0.;
            ay = 0.;
        }
        qx = ax * ca - ay * sa;
        qy = ax * sa + ay * ca;
    }
    if (EXTRA && o.runoff) {
        // swSp::sourceRunoffEverywhere / sourceRunoffWateronly
        // (swSpSourceRunoffInfiltration.cpp:43-263): rain rate of the cell's
        // station at t + dt/2; rain adds (a share of) the depth, evaporation
        // removes it and scales the discharge with the depth
        const double r = record_value(m.rain_t, m.rain_v + (size_t)m.rcode[i] * m.rain_nt,
                                      m.rain_nt, ck->rain_k, ck->t_half);
        const double h0 = h;
        double dp;
        if (o.rain_water_only) {
            dp = dt * r;
        } else {
            const double ph = m.phi[i];
            dp = 0.0;
            if (ph > kSmall) {
                dp = dt * r;
                if (o.runoff != SW2DG_RUNOFF_RUNOFF_COEFF || dp >= 0.0) dp /= ph;
            }
        }
        if (dp > 0.) {
            if (o.runoff == SW2DG_RUNOFF_FULL)
                h += dp;
            else if (o.runoff == SW2DG_RUNOFF_INFIL_COEFF)
                h += (1.0 - m.rcoef[i]) * dp;
            else
                h += m.rcoef[i] * dp;
        } else {
            h = smax(0.0, h0 + dp);
            const double ratio = h / smax(h0, kSmall);
            qx *= ratio;
            qy *= ratio;
        }
    }
    if (EXTRA && o.infil) {
        // swSp::sourceInfiltration, Horton (swSpSourceRunoffInfiltration.cpp:12-40)
        const double h0 = h;
        h = smax(0.0, h0 + -m.infil[i] * dt);
        const double ratio = h / smax(h0, kSmall);
        qx *= ratio;
        qy *= ratio;
    }
    if (EXTRA && o.wind) {
        // swSp::sourceWind + sourceWindSmithBanke (swSpSourceWind.cpp:11-85):
        // station velocity interpolated at t + dt/2, then the Smith & Banke
        // stress (its drag coefficient grows with the water speed |q|/h)
        const int k = ck->wind_k, nt = m.wind_nt;
        const double wx = record_value(m.wind_t, m.wind_x + (size_t)m.wcode[i] * nt, nt, k, ck->t_half);
        const double wy = record_value(m.wind_t, m.wind_y + (size_t)m.wcode[i] * nt, nt, k, ck->t_half);
        if (h > kSmall) {
            const double u = qx / h, v = qy / h;
            const double rx = wx - u, ry = wy - v;
            const double un = sqrt(u * u + v * v);
            const double rn = sqrt(rx * rx + ry * ry);
            const double csd = 1e-3 * (0.63 + 0.066 * un);
            const double mc = smin(h, o.h_min) / smax(o.h_min, kSmall);
            qx = qx + sp::kRhoAir * csd * rn * rx / sp::kRhoWater * dt * mc;
            qy = qy + sp::kRhoAir * csd * rn * ry / sp::kRhoWater * dt * mc;
        }
    }
    if (h > o.h_min) {
        const double qn = sqrt(qx * qx + qy * qy);
        const double c = sqrt(kG * h);
        const double fr = qn / c / h;
        const double fact = smax(qn / (h * o.v_max), fr / o.fr_max);
        if (fact > 1.0) {
            qx = qx / fact;
            qy = qy / fact;
        }
    }
    m.h[i] = h;
    m.qx[i] = qx;
    m.qy[i] = qy;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Hilbert index of (x, y) on a 2^16 x 2^16 grid (classic xy -> d walk)
uint32_t hilbert_d(uint32_t x, uint32_t y) {
    const uint32_t n = 1u << 16;
    uint32_t d = 0;
    for (uint32_t s = n >> 1; s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            const uint32_t tmp = x;
            x = y;
            y = tmp;
        }
    }
    return d;
}

}  // namespace

namespace sw2dg_detail {
int fail(sw2dg_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}
}  // namespace sw2dg_detail

namespace {

void launch_update(sw2dg_ctx *c, int p) {
    const dim3 blk(kBlock);
    if (!c->o.mdiff) {
        if (c->o.runoff || c->o.infil || c->o.wind)
            hipLaunchKernelGGL(k_update<true>, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->o, c->bc,
                               c->ck, p, 0);
        else
            hipLaunchKernelGGL(k_update<false>, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->o, c->bc,
                               c->ck, p, 0);
        return;
    }
    hipLaunchKernelGGL(k_update<true>, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->o, c->bc, c->ck, p, 1);
    hipLaunchKernelGGL(k_parab_dt, dim3(nblocks(c->ni)), blk, 0, c->stream, c->m, c->ck, p);
    hipLaunchKernelGGL(k_parab_n, dim3(1), dim3(kDtBlock), 0, c->stream, c->m, c->o, c->ck, p,
                       nblocks(c->ni));
    hipLaunchKernelGGL(k_parab_sub, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->ck, p, c->m.qx, c->m.qy,
                       c->m.pqx, c->m.pqy);
    hipLaunchKernelGGL(k_update<true>, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->o, c->bc, c->ck, p, 2);
}

void launch_step(sw2dg_ctx *c, int p) {
    const dim3 blk(kBlock);
    if (c->o.scheme == 1) {
        hipLaunchKernelGGL(k_slopes, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->ck, p);
        hipLaunchKernelGGL(k_edges<1>, dim3(nblocks(c->ni)), blk, 0, c->stream, c->m, c->o, c->ck, p);
    } else {
        hipLaunchKernelGGL(k_edges<0>, dim3(nblocks(c->ni)), blk, 0, c->stream, c->m, c->o, c->ck, p);
    }
    hipLaunchKernelGGL(k_cfl, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->o, c->ck, p);
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(kDtBlock), 0, c->stream, c->m, c->bc, c->ck, p, nblocks(c->nc));
    launch_update(c, p);
}

int build_graph(sw2dg_ctx *c) {
    if (c->gexec) return SW2DG_OK;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < kStepsPerGraph; ++s) launch_step(c, s & 1);
    HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph));
    HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    return SW2DG_OK;
}

int pull_clock(sw2dg_ctx *c) {
    HIPCHK(c, hipMemcpyAsync(&c->hck, c->ck, sizeof(Clock), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SW2DG_OK;
}

int push_clock(sw2dg_ctx *c) {
    HIPCHK(c, hipMemcpyAsync(c->ck, &c->hck, sizeof(Clock), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SW2DG_OK;
}

// The momentum-diffusion substeps a step left pending (k_parab_n: n > 1),
// then the rest of that step (k_update part 2); slot q of the clock holds
// the step's outcome (the live slot).  The substeps alternate between
// (qx, qy) and (pqx, pqy); part 2 reads the latter.  Afterwards the next
// step's activity is restored.
int finish_parab(sw2dg_ctx *c, int q) {
    Clock &k = c->hck;
    const int p = 1 - q;  // the kernels run as the step of parity p: dt = DT[q]
    const int n = k.parab_n;
    k.parab_pending = 0;
    k.ACT[p] = 1;
    int st = push_clock(c);
    if (st) return st;
    const dim3 blk(kBlock), gc(nblocks(c->nc));
    double *bx[2] = {c->m.qx, c->m.pqx}, *by[2] = {c->m.qy, c->m.pqy};
    for (int s = 0; s < n; ++s)
        hipLaunchKernelGGL(k_parab_sub, gc, blk, 0, c->stream, c->m, c->ck, p, bx[s & 1], by[s & 1],
                           bx[(s + 1) & 1], by[(s + 1) & 1]);
    if ((n & 1) == 0) {  // the result is in (qx, qy)
        HIPCHK(c, hipMemcpyAsync(c->m.pqx, c->m.qx, c->nc * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->m.pqy, c->m.qy, c->nc * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
    }
    hipLaunchKernelGGL(k_update<true>, gc, blk, 0, c->stream, c->m, c->o, c->bc, c->ck, p, 2);
    HIPCHK(c, hipGetLastError());
    if ((st = pull_clock(c))) return st;
    k.ACT[p] = 0;
    k.ACT[q] = k.act_saved;
    return push_clock(c);
}

int map_err(int st) {
    switch (st) {
        case sp::ST_NAN: return SW2DG_ERR_NAN;
        case sp::ST_NEG_DEPTH: return SW2DG_ERR_NEG_DEPTH;
        case sp::ST_FLOW_CONFIG: return SW2DG_ERR_FLOW_CONFIG;
        case sp::ST_BC: return SW2DG_ERR_BC;
        case sp::ST_PARAB_SUBSTEPS: return SW2DG_ERR_ARG;
        default: return SW2DG_ERR_HIP;
    }
}

const char *err_text(int st) {
    switch (st) {
        case sp::ST_NAN: return "SP - Mass flux is Not a number for an internal interface";
        case sp::ST_NEG_DEPTH: return "SP - Negative water depth";
        case sp::ST_FLOW_CONFIG: return "Impossible flow configuration in HLLC";
        case sp::ST_BC: return "invalid boundary condition value or type";
        case sp::ST_PARAB_SUBSTEPS:
            return "momentum diffusion needs more than 2^20 explicit substeps in one step";
        default: return "device error";
    }
}

}  // namespace

namespace sw2dg_detail {
// Renumbering and device-order tables of one mesh (shared by the single-run
// ctx and the ensemble): Hilbert order of the cell centres, internal edges by
// (lower, upper) renumbered cell, boundary edges by cell, incidence lists in
// the reference's order (swParent.cpp:469-486).
void build_tables(const sw2dg_mesh *mesh, MeshTables &T) {
    const int64_t nc = mesh->n_cell, nb = mesh->n_bnd, ni = mesh->n_int;
    T.nc = nc;
    T.nb = nb;
    T.ni = ni;
    // ---- Hilbert renumbering of cells
    double xmin = mesh->cell_x[0], xmax = xmin, ymin = mesh->cell_y[0], ymax = ymin;
    for (int64_t i = 1; i < nc; ++i) {
        xmin = std::min(xmin, mesh->cell_x[i]);
        xmax = std::max(xmax, mesh->cell_x[i]);
        ymin = std::min(ymin, mesh->cell_y[i]);
        ymax = std::max(ymax, mesh->cell_y[i]);
    }
    const double sc = 65535.0 / std::max({xmax - xmin, ymax - ymin, 1e-30});
    std::vector<uint32_t> key(nc);
    for (int64_t i = 0; i < nc; ++i) {
        const uint32_t hx = (uint32_t)std::min(65535.0, std::max(0.0, (mesh->cell_x[i] - xmin) * sc));
        const uint32_t hy = (uint32_t)std::min(65535.0, std::max(0.0, (mesh->cell_y[i] - ymin) * sc));
        key[i] = hilbert_d(hx, hy);
    }
    T.perm.resize(nc);
    std::iota(T.perm.begin(), T.perm.end(), 0);
    std::stable_sort(T.perm.begin(), T.perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    T.inv.resize(nc);
    for (int64_t k = 0; k < nc; ++k) T.inv[T.perm[k]] = (int)k;
    // ---- internal edges sorted by (lower, upper) renumbered cell
    T.eperm.resize(ni);
    std::iota(T.eperm.begin(), T.eperm.end(), 0);
    std::vector<int64_t> ekey(ni);
    for (int64_t e = 0; e < ni; ++e) {
        const int64_t a = T.inv[mesh->int_L[e]], b = T.inv[mesh->int_R[e]];
        ekey[e] = std::min(a, b) * (int64_t)nc + std::max(a, b);
    }
    std::stable_sort(T.eperm.begin(), T.eperm.end(), [&](int a, int b) { return ekey[a] < ekey[b]; });
    std::vector<int> einv(ni);
    for (int64_t k = 0; k < ni; ++k) einv[T.eperm[k]] = (int)k;
    // ---- boundary edges sorted by renumbered cell (stable: keeps reference order)
    T.bperm.resize(nb);
    std::iota(T.bperm.begin(), T.bperm.end(), 0);
    std::stable_sort(T.bperm.begin(), T.bperm.end(),
                     [&](int a, int b) { return T.inv[mesh->bnd_cell[a]] < T.inv[mesh->bnd_cell[b]]; });
    T.binv.resize(nb);
    for (int64_t k = 0; k < nb; ++k) T.binv[T.bperm[k]] = (int)k;
    // ---- incidence lists in reference order (swParent.cpp:469-486)
    auto &iptr = T.iptr, &bptr = T.bptr;
    iptr.assign(nc + 1, 0);
    bptr.assign(nc + 1, 0);
    for (int64_t e = 0; e < ni; ++e) {
        iptr[T.inv[mesh->int_L[e]] + 1]++;
        iptr[T.inv[mesh->int_R[e]] + 1]++;
    }
    for (int64_t e = 0; e < nb; ++e) bptr[T.inv[mesh->bnd_cell[e]] + 1]++;
    for (int64_t i = 0; i < nc; ++i) {
        iptr[i + 1] += iptr[i];
        bptr[i + 1] += bptr[i];
    }
    auto &iinc = T.iinc, &binc = T.binc;
    iinc.assign(2 * ni, 0);
    binc.assign(nb, 0);
    std::vector<int> fill_i(iptr.begin(), iptr.end() - 1), fill_b(bptr.begin(), bptr.end() - 1);
    for (int64_t e = 0; e < ni; ++e) {  // ascending reference edge index
        const int ne = einv[e];
        iinc[fill_i[T.inv[mesh->int_L[e]]]++] = (ne << 1) | 0;
        iinc[fill_i[T.inv[mesh->int_R[e]]]++] = (ne << 1) | 1;
    }
    for (int64_t e = 0; e < nb; ++e) binc[fill_b[T.inv[mesh->bnd_cell[e]]]++] = T.binv[e];
    // ---- gather arrays into the new order
    auto &eL = T.eL, &eR = T.eR;
    auto &elen = T.elen, &enx = T.enx, &eny = T.eny, &exm = T.exm, &eym = T.eym;
    eL.assign(ni, 0);
    eR.assign(ni, 0);
    elen.assign(ni, 0.0);
    enx.assign(ni, 0.0);
    eny.assign(ni, 0.0);
    exm.assign(ni, 0.0);
    eym.assign(ni, 0.0);
    for (int64_t k = 0; k < ni; ++k) {
        const int e = T.eperm[k];
        eL[k] = T.inv[mesh->int_L[e]];
        eR[k] = T.inv[mesh->int_R[e]];
        elen[k] = mesh->int_len[e];
        enx[k] = mesh->int_nx[e];
        eny[k] = mesh->int_ny[e];
        exm[k] = mesh->int_xm[e];
        eym[k] = mesh->int_ym[e];
    }
    auto &blen = T.blen, &bnx = T.bnx, &bny = T.bny;
    auto &bser = T.bser, &bcell = T.bcell, &bcells = T.bcells;
    blen.assign(nb, 0.0);
    bnx.assign(nb, 0.0);
    bny.assign(nb, 0.0);
    bser.assign(nb, 0);
    bcell.assign(std::max<int64_t>(nb, 1), 0);
    bcells.clear();
    for (int64_t k = 0; k < nb; ++k) {
        const int e = T.bperm[k];
        blen[k] = mesh->bnd_len[e];
        bnx[k] = mesh->bnd_nx[e];
        bny[k] = mesh->bnd_ny[e];
        bser[k] = mesh->bnd_series[e];
        bcell[k] = T.inv[mesh->bnd_cell[e]];
    }
    for (int64_t i = 0; i < nc; ++i)
        if (bptr[i + 1] > bptr[i]) bcells.push_back((int)i);
    T.n_bcells = (int)bcells.size();
    if (bcells.empty()) bcells.push_back(0);
}

}  // namespace sw2dg_detail

extern "C" {

const char *sw2dg_version(void) { return "sw2dg 0.1 (gfx950, FP64, fp-contract off)"; }

int sw2dg_ctx_create(const sw2dg_mesh *mesh, const sw2dg_params *prm, int32_t device,
                     sw2dg_ctx **out) {
    if (!out) return SW2DG_ERR_ARG;
    *out = nullptr;
    if (!mesh || !prm) return SW2DG_ERR_ARG;
    const int64_t nc = mesh->n_cell, nb = mesh->n_bnd, ni = mesh->n_int;
    if (nc <= 0 || nb < 0 || ni < 0 || nc > INT32_MAX / 2 || ni > INT32_MAX / 8)
        return SW2DG_ERR_ARG;
    if (!mesh->cell_area || !mesh->cell_x || !mesh->cell_y || !mesh->cell_zb) return SW2DG_ERR_ARG;
    if (nb > 0 && (!mesh->bnd_cell || !mesh->bnd_len || !mesh->bnd_nx || !mesh->bnd_ny ||
                   !mesh->bnd_series))
        return SW2DG_ERR_ARG;
    if (ni > 0 && (!mesh->int_L || !mesh->int_R || !mesh->int_len || !mesh->int_nx ||
                   !mesh->int_ny || !mesh->int_xm || !mesh->int_ym))
        return SW2DG_ERR_ARG;
    for (int64_t e = 0; e < nb; ++e)
        if (mesh->bnd_cell[e] < 0 || mesh->bnd_cell[e] >= nc || mesh->bnd_series[e] < 0)
            return SW2DG_ERR_ARG;
    for (int64_t e = 0; e < ni; ++e)
        if (mesh->int_L[e] < 0 || mesh->int_L[e] >= nc || mesh->int_R[e] < 0 ||
            mesh->int_R[e] >= nc || mesh->int_L[e] == mesh->int_R[e])
            return SW2DG_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SW2DG_ERR_NO_DEVICE;
    sw2dg_ctx *c = new sw2dg_ctx();
    if (device < 0) {
        hipGetDevice(&c->device);
    } else {
        c->device = device;
    }
    if (hipSetDevice(c->device) != hipSuccess) {
        delete c;
        return SW2DG_ERR_NO_DEVICE;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SW2DG_ERR_HIP;
    }
    c->prm = *prm;
    c->nc = nc;
    c->nb = nb;
    c->ni = ni;
    MeshTables T;
    build_tables(mesh, T);
    c->perm = T.perm;
    c->inv = T.inv;
    c->eperm = T.eperm;
    c->bperm = T.bperm;
    c->binv = T.binv;
    auto &iptr = T.iptr, &bptr = T.bptr, &iinc = T.iinc, &binc = T.binc;
    auto &eL = T.eL, &eR = T.eR;
    auto &elen = T.elen, &enx = T.enx, &eny = T.eny, &exm = T.exm, &eym = T.eym;
    auto &blen = T.blen, &bnx = T.bnx, &bny = T.bny;
    auto &bser = T.bser, &bcell = T.bcell, &bcells = T.bcells;
    const int n_bcells = T.n_bcells;
    auto cellv = [&](const double *src) {
        std::vector<double> v(nc);
        for (int64_t k = 0; k < nc; ++k) v[k] = src[c->perm[k]];
        return v;
    };
    c->bser_host.assign(mesh->bnd_series, mesh->bnd_series + nb);
    MeshDev &m = c->m;
    m.n_cell = (int)nc;
    m.n_bnd = (int)nb;
    m.n_int = (int)ni;
    m.h = c->dalloc<double>(nc);
    m.qx = c->dalloc<double>(nc);
    m.qy = c->dalloc<double>(nc);
    m.zb = c->upload(cellv(mesh->cell_zb));
    m.area = c->upload(cellv(mesh->cell_area));
    m.cx = c->upload(cellv(mesh->cell_x));
    m.cy = c->upload(cellv(mesh->cell_y));
    std::vector<double> ones(nc, 1.0), zeros(nc, 0.0);
    m.phi = c->upload(ones);
    m.beta = c->upload(ones);
    m.fc1 = c->upload(zeros);
    m.fc2 = c->upload(zeros);
    m.fcos = c->upload(ones);
    m.fsin = c->upload(zeros);
    m.hsing = c->upload(zeros);
    m.hk1 = c->upload(zeros);
    m.hk2 = c->upload(zeros);
    m.hcos = c->upload(ones);
    m.hsin = c->upload(zeros);
    m.dzdx = c->dalloc<double>(nc);
    m.dzdy = c->dalloc<double>(nc);
    m.int_ptr = c->upload(iptr);
    m.int_inc = c->upload(iinc);
    {
        std::vector<int> inbr(iinc.size());
        for (size_t k = 0; k < iinc.size(); ++k) inbr[k] = (iinc[k] & 1) ? eL[iinc[k] >> 1] : eR[iinc[k] >> 1];
        m.int_nbr = c->upload(inbr);
    }
    m.bnd_ptr = c->upload(bptr);
    m.bnd_inc = c->upload(binc);
    m.eL = c->upload(eL);
    m.eR = c->upload(eR);
    m.elen = c->upload(elen);
    m.enx = c->upload(enx);
    m.eny = c->upload(eny);
    m.exm = c->upload(exm);
    m.eym = c->upload(eym);
    m.flx = c->dalloc<double>(6 * (size_t)ni);
    m.swp = c->dalloc<double>(2 * (size_t)ni);
    m.blen = c->upload(blen);
    m.bnx = c->upload(bnx);
    m.bny = c->upload(bny);
    m.bser = c->upload(bser);
    m.bcell = c->upload(bcell);
    m.bnd_cells = c->upload(bcells);
    m.n_bnd_cells = n_bcells;
    m.bF1 = c->dalloc<double>(nb);
    m.partial = c->dalloc<double>(nblocks(nc));
    c->ck = c->dalloc<Clock>(1);
    c->o.scheme = prm->scheme;
    c->o.topo = prm->topo;
    c->o.friction = prm->friction;
    c->o.head_loss = prm->head_loss;
    c->o.neg_det = prm->neg_depth_detection;
    c->o.h_min = prm->h_min;
    c->o.phi_min = prm->phi_min;
    c->o.v_max = prm->v_max;
    c->o.fr_max = prm->fr_max;
    c->o.adj_cut = 0.0;
    // a default "no series" BC table (every boundary edge must then have series 0 set later)
    for (void *p : c->allocs)
        if (!p) {
            sw2dg_ctx_destroy(c);
            return SW2DG_ERR_HIP;
        }
    std::memset(&c->hck, 0, sizeof(Clock));
    c->hck.t_max = prm->t_max;
    c->hck.dt_max = prm->dt_max;
    c->hck.dt_map = prm->dt_map;
    c->hck.dt_ts = prm->dt_time_series;
    c->hck.cfl = prm->cfl_max;
    c->hck.T[0] = prm->t0;
    *out = c;
    return SW2DG_OK;
}

int sw2dg_set_cell_params(sw2dg_ctx *c, const double *phi, const double *beta,
                          const double *fric_c1, const double *fric_c2,
                          const double *fric_alpha_deg, const double *hl_hsing,
                          const double *hl_c1, const double *hl_c2, const double *hl_alpha_deg) {
    if (!c) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nc = c->nc;
    auto put = [&](const double *src, const double *dst) -> int {
        std::vector<double> v(nc);
        for (int64_t k = 0; k < nc; ++k) v[k] = src[c->perm[k]];
        HIPCHK(c, hipMemcpy((void *)dst, v.data(), nc * sizeof(double), hipMemcpyHostToDevice));
        return SW2DG_OK;
    };
    int st;
    if (phi && (st = put(phi, c->m.phi))) return st;
    if (beta && (st = put(beta, c->m.beta))) return st;
    if (fric_c1 && (st = put(fric_c1, c->m.fc1))) return st;
    if (fric_c2 && (st = put(fric_c2, c->m.fc2))) return st;
    if (fric_alpha_deg) {
        // degrees -> radians and cos/sin on the host, as swSp.cpp:1066-1069
        std::vector<double> cs(nc), sn(nc);
        for (int64_t i = 0; i < nc; ++i) {
            const double a = fric_alpha_deg[i] * atan(1.0) / 45.0;
            cs[i] = cos(a);
            sn[i] = sin(a);
        }
        if ((st = put(cs.data(), c->m.fcos))) return st;
        if ((st = put(sn.data(), c->m.fsin))) return st;
    }
    if (hl_hsing && (st = put(hl_hsing, c->m.hsing))) return st;
    if (hl_c1 && (st = put(hl_c1, c->m.hk1))) return st;
    if (hl_c2 && (st = put(hl_c2, c->m.hk2))) return st;
    if (hl_alpha_deg) {
        // swSp.cpp:1365 then cos/sin as swSpSourceFriction.cpp:186-188
        std::vector<double> cs(nc), sn(nc);
        for (int64_t i = 0; i < nc; ++i) {
            const double a = hl_alpha_deg[i] * atan(1.0) / 45.0;
            cs[i] = cos(a);
            sn[i] = sin(a);
        }
        if ((st = put(cs.data(), c->m.hcos))) return st;
        if ((st = put(sn.data(), c->m.hsin))) return st;
    }
    return SW2DG_OK;
}

int sw2dg_set_bc_series(sw2dg_ctx *c, int32_t n_series, int32_t n_times, const int32_t *type,
                        const double *t, const double *v) {
    if (!c || n_series < 0 || (n_series > 0 && (n_times < 1 || !type || !t || !v)))
        return SW2DG_ERR_ARG;
    for (int k = 0; k < n_series; ++k)
        if (type[k] < 0 || type[k] > 3) return fail(c, SW2DG_ERR_BC, "BC type not handled");
    for (int64_t e = 0; e < c->nb; ++e)
        if (c->bser_host[e] >= n_series)
            return fail(c, SW2DG_ERR_ARG, "a boundary edge refers to a missing BC series");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->bc.t && n_series == c->n_series && n_times == c->n_times && n_series > 0) {
        // same shape (a new hydrograph in a calibration loop): overwrite in place
        HIPCHK(c, hipMemcpy((void *)c->bc.t, t, n_times * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->bc.v, v, (size_t)n_series * n_times * sizeof(double),
                            hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->bc.type, type, n_series * sizeof(int), hipMemcpyHostToDevice));
        c->bc.sorted = std::is_sorted(t, t + n_times) ? 1 : 0;
        c->series_t_host.assign(t, t + n_times);
        c->series_v_host.assign(v, v + (size_t)n_series * n_times);
        c->series_type_host.assign(type, type + n_series);
        return SW2DG_OK;
    }
    c->n_series = n_series;
    c->n_times = n_times;
    const int ns = std::max(n_series, 1), nt = std::max(n_times, 1);
    std::vector<double> tv(t, t + n_times), vv(v, v + (size_t)n_series * n_times);
    std::vector<int> ty(type, type + n_series);
    if (tv.empty()) tv.push_back(0.0);
    if (vv.empty()) vv.push_back(0.0);
    if (ty.empty()) ty.push_back(3);
    c->series_t_host = tv;
    c->series_v_host = vv;
    c->series_type_host = ty;
    c->bc.t = c->upload(tv);
    c->bc.v = c->upload(vv);
    c->bc.type = c->upload(ty);
    c->bc.vinf = c->dalloc<double>(ns);
    c->bc.vsup = c->dalloc<double>(ns);
    c->bc.tinf = c->dalloc<double>(ns);
    c->bc.tsup = c->dalloc<double>(ns);
    c->bc.val = c->dalloc<double>(ns);
    c->bc.rank = c->dalloc<int>(ns);
    c->bc.atend = c->dalloc<int>(ns);
    c->bc.n_times = nt;
    c->bc.sorted = std::is_sorted(tv.begin(), tv.end()) ? 1 : 0;
    c->hck.n_series = n_series;
    if (c->gexec) {  // arguments changed: rebuild the graph
        hipGraphExecDestroy(c->gexec);
        hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
    }
    return SW2DG_OK;
}

int sw2dg_set_wind(sw2dg_ctx *c, int32_t n_station, int32_t n_times, const int32_t *code,
                   const double *t, const double *vx, const double *vy) {
    if (!c || n_station < 0) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (n_station > 0 && c->assim)
        return fail(c, SW2DG_ERR_ARG, "the 4D-Var layer has no wind source");
    if (n_station == 0) {
        c->o.wind = 0;
    } else {
        if (n_times < 1 || !code || !t || !vx || !vy) return SW2DG_ERR_ARG;
        for (int k = 1; k < n_times; ++k)
            if (t[k - 1] > t[k])
                return fail(c, SW2DG_ERR_ARG, "wind record times must be non-decreasing");
        std::vector<int> wc(c->nc);
        for (int64_t k = 0; k < c->nc; ++k) {
            const int32_t s = code[c->perm[k]];
            if (s < 1 || s > n_station)
                return fail(c, SW2DG_ERR_ARG, "a wind station code lies outside [1, n_station]");
            wc[k] = s - 1;
        }
        const size_t nv = (size_t)n_station * n_times;
        if (!c->m.wcode) c->m.wcode = c->upload(wc);
        else HIPCHK(c, hipMemcpy((void *)c->m.wcode, wc.data(), wc.size() * sizeof(int),
                                 hipMemcpyHostToDevice));
        if (c->m.wind_nt != n_times || c->wind_nv != nv) {
            // a new table shape: fresh buffers (the old ones stay owned by the ctx)
            c->m.wind_t = c->dalloc<double>(n_times);
            c->m.wind_x = c->dalloc<double>(nv);
            c->m.wind_y = c->dalloc<double>(nv);
            c->wind_nv = nv;
        }
        if (!c->m.wcode || !c->m.wind_t || !c->m.wind_x || !c->m.wind_y)
            return fail(c, SW2DG_ERR_HIP, "wind tables: device allocation failed");
        HIPCHK(c, hipMemcpy((void *)c->m.wind_t, t, n_times * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->m.wind_x, vx, nv * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->m.wind_y, vy, nv * sizeof(double), hipMemcpyHostToDevice));
        c->o.wind = 1;
    }
    c->m.wind_nt = c->o.wind ? n_times : 0;
    if (c->gexec) {  // kernel arguments changed: rebuild the graph
        hipGraphExecDestroy(c->gexec);
        hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
    }
    return SW2DG_OK;
}

int sw2dg_set_runoff(sw2dg_ctx *c, int32_t model, int32_t water_only, int32_t n_station,
                     int32_t n_times, const int32_t *code, const double *t, const double *rate,
                     const double *coeff) {
    if (!c || model < SW2DG_RUNOFF_NONE || model > SW2DG_RUNOFF_RUNOFF_COEFF) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (model != SW2DG_RUNOFF_NONE && c->assim)
        return fail(c, SW2DG_ERR_ARG, "the 4D-Var layer has no runoff source");
    if (model != SW2DG_RUNOFF_NONE) {
        if (n_station < 1 || n_times < 1 || !code || !t || !rate ||
            (model != SW2DG_RUNOFF_FULL && !coeff))
            return SW2DG_ERR_ARG;
        for (int k = 1; k < n_times; ++k)
            if (t[k - 1] > t[k])
                return fail(c, SW2DG_ERR_ARG, "rain record times must be non-decreasing");
        std::vector<int> rc(c->nc);
        std::vector<double> cf(c->nc, 0.0);
        for (int64_t k = 0; k < c->nc; ++k) {
            const int32_t s = code[c->perm[k]];
            if (s < 1 || s > n_station)
                return fail(c, SW2DG_ERR_ARG, "a rain station code lies outside [1, n_station]");
            rc[k] = s - 1;
            if (model != SW2DG_RUNOFF_FULL) {
                cf[k] = coeff[c->perm[k]];
                if (!(cf[k] >= 0.0 && cf[k] <= 1.0))
                    return fail(c, SW2DG_ERR_ARG, "a runoff coefficient lies outside [0, 1]");
            }
        }
        const size_t nv = (size_t)n_station * n_times;
        if (!c->m.rcode) {
            c->m.rcode = c->dalloc<int>(c->nc);
            c->m.rcoef = c->dalloc<double>(c->nc);
        }
        if (c->m.rain_nt != n_times || c->rain_nv != nv) {
            c->m.rain_t = c->dalloc<double>(n_times);
            c->m.rain_v = c->dalloc<double>(nv);
            c->rain_nv = nv;
        }
        if (!c->m.rcode || !c->m.rcoef || !c->m.rain_t || !c->m.rain_v)
            return fail(c, SW2DG_ERR_HIP, "rain tables: device allocation failed");
        HIPCHK(c, hipMemcpy((void *)c->m.rcode, rc.data(), rc.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->m.rcoef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->m.rain_t, t, n_times * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((void *)c->m.rain_v, rate, nv * sizeof(double), hipMemcpyHostToDevice));
    }
    c->o.runoff = model;
    c->o.rain_water_only = water_only ? 1 : 0;
    c->m.rain_nt = model != SW2DG_RUNOFF_NONE ? n_times : 0;
    if (c->gexec) {  // kernel arguments changed: rebuild the graph
        hipGraphExecDestroy(c->gexec);
        hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
    }
    return SW2DG_OK;
}

int sw2dg_set_infiltration(sw2dg_ctx *c, int32_t model, const double *rate) {
    if (!c || (model != SW2DG_INFIL_NONE && model != SW2DG_INFIL_HORTON)) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (model != SW2DG_INFIL_NONE) {
        if (!rate) return SW2DG_ERR_ARG;
        if (c->assim) return fail(c, SW2DG_ERR_ARG, "the 4D-Var layer has no infiltration source");
        std::vector<double> v(c->nc);
        for (int64_t k = 0; k < c->nc; ++k) {
            v[k] = rate[c->perm[k]];
            if (!(v[k] >= 0.0)) return fail(c, SW2DG_ERR_ARG, "an infiltration rate is negative");
        }
        if (!c->m.infil) c->m.infil = c->dalloc<double>(c->nc);
        if (!c->m.infil) return fail(c, SW2DG_ERR_HIP, "infiltration map: device allocation failed");
        HIPCHK(c, hipMemcpy((void *)c->m.infil, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c->o.infil = model;
    if (c->gexec) {
        hipGraphExecDestroy(c->gexec);
        hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
    }
    return SW2DG_OK;
}

int sw2dg_set_momentum_diffusion(sw2dg_ctx *c, int32_t model) {
    if (!c || model < 0 || model > 1) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (model && c->assim) return fail(c, SW2DG_ERR_ARG, "the 4D-Var layer has no momentum diffusion");
    if (model && !c->m.pfac) {
        double *fac = c->dalloc<double>(c->ni);
        c->m.pqx = c->dalloc<double>(c->nc);
        c->m.pqy = c->dalloc<double>(c->nc);
        c->m.ppart = c->dalloc<double>(nblocks(c->ni));
        if (!fac || !c->m.pqx || !c->m.pqy || !c->m.ppart)
            return fail(c, SW2DG_ERR_HIP, "momentum diffusion: device allocation failed");
        if (c->ni > 0)
            hipLaunchKernelGGL(k_parab_fac, dim3(nblocks(c->ni)), dim3(kBlock), 0, c->stream, c->m, fac);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->m.pfac = fac;
    }
    c->o.mdiff = model;
    c->o.parab_span = c->prm.t_max - c->prm.t0;
    if (c->gexec) {
        hipGraphExecDestroy(c->gexec);
        hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
    }
    return SW2DG_OK;
}

int sw2dg_set_state(sw2dg_ctx *c, const double *h, const double *qx, const double *qy, double t) {
    if (!c || !h || !qx || !qy) return SW2DG_ERR_ARG;
    if (!c->bc.t) return fail(c, SW2DG_ERR_ARG, "sw2dg_set_bc_series must be called first");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nc = c->nc;
    std::vector<double> v(nc);
    double *dst[3] = {c->m.h, c->m.qx, c->m.qy};
    const double *src[3] = {h, qx, qy};
    std::vector<double> *keep[3] = {&c->h0, &c->qx0, &c->qy0};
    for (int a = 0; a < 3; ++a) {
        for (int64_t k = 0; k < nc; ++k) v[k] = src[a][c->perm[k]];
        HIPCHK(c, hipMemcpy(dst[a], v.data(), nc * sizeof(double), hipMemcpyHostToDevice));
        *keep[a] = v;
    }
    c->t0 = t;
    // fluxes of dry-dry edges start at zero (reference vectors are value-initialised)
    HIPCHK(c, hipMemset(c->m.flx, 0, 6 * (size_t)std::max<int64_t>(c->ni, 1) * sizeof(double)));
    HIPCHK(c, hipMemset(c->m.swp, 0, 2 * (size_t)std::max<int64_t>(c->ni, 1) * sizeof(double)));
    // BC cursors rewound (swBoundaryConditionTimeSeries::initialise)
    const int ns = c->n_series;
    if (ns > 0) {
        std::vector<double> tv(c->n_times), vinf(ns), vsup(ns), tinf(ns), tsup(ns);
        std::vector<int> rank(ns, 1), atend(ns, 0);
        std::vector<double> vv((size_t)ns * c->n_times);
        HIPCHK(c, hipMemcpy(tv.data(), c->bc.t, tv.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(vv.data(), c->bc.v, vv.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < ns; ++k) {
            vinf[k] = vv[(size_t)k * c->n_times];
            tinf[k] = tv[0];
            if (c->n_times > 1) {
                vsup[k] = vv[(size_t)k * c->n_times + 1];
                tsup[k] = tv[1];
            } else {
                vsup[k] = vinf[k];
                tsup[k] = tinf[k];
                atend[k] = 1;
            }
        }
        HIPCHK(c, hipMemcpy(c->bc.vinf, vinf.data(), ns * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->bc.vsup, vsup.data(), ns * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->bc.tinf, tinf.data(), ns * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->bc.tsup, tsup.data(), ns * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->bc.rank, rank.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->bc.atend, atend.data(), ns * sizeof(int), hipMemcpyHostToDevice));
    }
    c->hck.T[0] = t;
    c->hck.T[1] = t;
    c->hck.DT[0] = 0.0;
    c->hck.DT[1] = 0.0;
    c->hck.STEP[0] = c->hck.STEP[1] = 0;
    c->hck.ACT[0] = c->hck.ACT[1] = 0;
    c->hck.err = 0;
    c->hck.paused = 0;
    c->hck.store_flags = 0;
    c->hck.dt_last = 0.0;
    c->hck.parab_pending = 0;
    c->state_set = true;
    return push_clock(c);
}

int sw2dg_get_state(sw2dg_ctx *c, double *h, double *qx, double *qy) {
    if (!c) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t nc = c->nc;
    std::vector<double> v(nc);
    const double *srcs[3] = {c->m.h, c->m.qx, c->m.qy};
    double *dsts[3] = {h, qx, qy};
    for (int a = 0; a < 3; ++a) {
        if (!dsts[a]) continue;
        HIPCHK(c, hipMemcpy(v.data(), srcs[a], nc * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < nc; ++k) dsts[a][c->perm[k]] = v[k];
    }
    return SW2DG_OK;
}

int sw2dg_run(sw2dg_ctx *c, double t_stop, int64_t max_steps, int32_t pause_on_store,
              int64_t *n_done) {
    if (n_done) *n_done = 0;
    if (!c || max_steps < 0) return SW2DG_ERR_ARG;
    if (!c->state_set) return fail(c, SW2DG_ERR_ARG, "sw2dg_set_state must be called first");
    HIPCHK(c, hipSetDevice(c->device));
    int st = pull_clock(c);
    if (st) return st;
    if (c->hck.err) return fail(c, map_err(c->hck.err), err_text(c->hck.err));
    const long long step0 = c->hck.STEP[0];
    c->hck.t_stop = t_stop;
    c->hck.step_stop = step0 + max_steps;
    c->hck.pause_on_store = pause_on_store;
    c->hck.paused = 0;
    const double t = c->hck.T[0];
    c->hck.ACT[0] = (t < c->hck.t_max) && (t < t_stop) && (max_steps > 0);
    c->hck.T[1] = t;
    c->hck.DT[1] = c->hck.DT[0];
    c->hck.STEP[1] = step0;
    c->hck.ACT[1] = 0;
    if ((st = push_clock(c))) return st;
    if ((st = build_graph(c))) return st;
    while (c->hck.ACT[0]) {
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        if ((st = pull_clock(c))) return st;
        if (c->hck.err) break;
        // a step with more than one diffusion substep idled the rest of the
        // launch: the live clock is in slot 0 (the idle steps carried it)
        if (c->hck.parab_pending && (st = finish_parab(c, 0))) return st;
        if (c->hck.err) break;
    }
    if (n_done) *n_done = c->hck.STEP[0] - step0;
    if (c->hck.err) return fail(c, map_err(c->hck.err), err_text(c->hck.err));
    return SW2DG_OK;
}

int sw2dg_get_clock(sw2dg_ctx *c, double *t, double *dt, int64_t *steps, int32_t *store_flags) {
    if (!c) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int st = pull_clock(c);
    if (st) return st;
    if (t) *t = c->hck.T[0];
    if (dt) *dt = c->hck.dt_last;
    if (steps) *steps = c->hck.STEP[0];
    if (store_flags) *store_flags = c->hck.store_flags;
    return SW2DG_OK;
}

int sw2dg_get_bnd_f1(sw2dg_ctx *c, double *f1) {
    if (!c || (!f1 && c->nb > 0)) return SW2DG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> v(c->nb);
    if (c->nb) HIPCHK(c, hipMemcpy(v.data(), c->m.bF1, c->nb * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < c->nb; ++k) f1[c->bperm[k]] = v[k];
    return SW2DG_OK;
}

int sw2dg_profile_kernels(sw2dg_ctx *c, int64_t n_steps, int32_t k_max, double *avg_ms,
                          const char **names, int32_t *n_kernels) {
    if (!c || n_steps <= 0 || !avg_ms || k_max <= 0) return SW2DG_ERR_ARG;
    if (!c->state_set) return fail(c, SW2DG_ERR_ARG, "sw2dg_set_state must be called first");
    HIPCHK(c, hipSetDevice(c->device));
    int st = pull_clock(c);
    if (st) return st;
    // run each step's kernels one by one, bracketed by events on the ctx stream
    const bool muscl = c->o.scheme == 1;
    const int nk = muscl ? 5 : 4;
    static const char *kn_m[5] = {"k_slopes", "k_edges<1>", "k_cfl", "k_dt", "k_update"};
    static const char *kn_g[4] = {"k_edges<0>", "k_cfl", "k_dt", "k_update"};
    std::vector<hipEvent_t> ev(2 * nk);
    for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
    std::vector<double> acc(nk, 0.0);
    int64_t done = 0;
    c->hck.t_stop = 1e300;
    c->hck.step_stop = c->hck.STEP[0] + n_steps;
    c->hck.pause_on_store = 0;
    c->hck.paused = 0;
    c->hck.ACT[0] = c->hck.T[0] < c->hck.t_max;
    if ((st = push_clock(c))) return st;
    const dim3 blk(kBlock);
    for (int64_t s = 0; s < n_steps; ++s) {
        const int p = (int)(s & 1);
        int k = 0;
        if (muscl) {
            hipEventRecord(ev[2 * k], c->stream);
            hipLaunchKernelGGL(k_slopes, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->ck, p);
            hipEventRecord(ev[2 * k + 1], c->stream);
            ++k;
            hipEventRecord(ev[2 * k], c->stream);
            hipLaunchKernelGGL(k_edges<1>, dim3(nblocks(c->ni)), blk, 0, c->stream, c->m, c->o, c->ck, p);
            hipEventRecord(ev[2 * k + 1], c->stream);
            ++k;
        } else {
            hipEventRecord(ev[2 * k], c->stream);
            hipLaunchKernelGGL(k_edges<0>, dim3(nblocks(c->ni)), blk, 0, c->stream, c->m, c->o, c->ck, p);
            hipEventRecord(ev[2 * k + 1], c->stream);
            ++k;
        }
        hipEventRecord(ev[2 * k], c->stream);
        hipLaunchKernelGGL(k_cfl, dim3(nblocks(c->nc)), blk, 0, c->stream, c->m, c->o, c->ck, p);
        hipEventRecord(ev[2 * k + 1], c->stream);
        ++k;
        hipEventRecord(ev[2 * k], c->stream);
        hipLaunchKernelGGL(k_dt, dim3(1), dim3(kDtBlock), 0, c->stream, c->m, c->bc, c->ck, p, nblocks(c->nc));
        hipEventRecord(ev[2 * k + 1], c->stream);
        ++k;
        hipEventRecord(ev[2 * k], c->stream);
        launch_update(c, p);
        hipEventRecord(ev[2 * k + 1], c->stream);
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->o.mdiff) {  // substeps left pending by the step (outside the timed kernels)
            if ((st = pull_clock(c))) return st;
            if (c->hck.parab_pending && (st = finish_parab(c, 1 - p))) return st;
        }
        for (int j = 0; j < nk; ++j) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev[2 * j], ev[2 * j + 1]);
            acc[j] += ms;
        }
        ++done;
    }
    if (n_steps & 1) {
        // an odd count leaves the live clock in slot 1: copy it back to slot 0
        if ((st = pull_clock(c))) return st;
        c->hck.T[0] = c->hck.T[1];
        c->hck.DT[0] = c->hck.DT[1];
        c->hck.STEP[0] = c->hck.STEP[1];
        c->hck.ACT[0] = c->hck.ACT[1];
        if ((st = push_clock(c))) return st;
    }
    for (auto &e : ev) hipEventDestroy(e);
    const int nout = std::min<int>(nk, k_max);
    for (int j = 0; j < nout; ++j) {
        avg_ms[j] = acc[j] / (double)done;
        if (names) names[j] = muscl ? kn_m[j] : kn_g[j];
    }
    if (n_kernels) *n_kernels = nk;
    if ((st = pull_clock(c))) return st;
    if (c->hck.err) return fail(c, map_err(c->hck.err), err_text(c->hck.err));
    return SW2DG_OK;
}

int64_t sw2dg_device_bytes(sw2dg_ctx *c) { return c ? c->bytes : 0; }

const char *sw2dg_last_error(sw2dg_ctx *c) { return c ? c->err.c_str() : "null context"; }

void sw2dg_ctx_destroy(sw2dg_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    sw2dg_detail::assim_destroy(c->assim);
    c->assim = nullptr;
    for (void *p : c->allocs)
        if (p) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
