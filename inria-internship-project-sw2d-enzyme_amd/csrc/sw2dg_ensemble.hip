// sw2dg_ensemble.hip -- batched ensemble of differentiable-forward (_E) runs
// over uneven meshes, in one persistent launch.
//
// Every run (its own mesh, controls, hydrograph, clock) is stepped to t_max
// by a cluster of c >= 1 workgroups: edge fluxes -> CFL minimum with the
// observation landing of build_checkpoint_records
// (assim/swSpE_checkpoint.cpp:224-366) -> cell update -> gauge sampling.
// The per-step operations are those of the single-run dry run (edge fluxes /
// k_e_cfl / k_e_dt / k_e_update in sw2dg_adjoint.hip), element by element
// and in the same order, so each run reproduces sw2dg_assim_record on its
// mesh bit for bit whatever its cluster size (the CFL minimum is exact in
// any order; every workgroup of a cluster derives the same dt from the same
// partial minima).
//
// Two launch modes:
//  * more runs than co-resident workgroups: one workgroup per run (c = 1),
//    phases separated by workgroup barriers, runs dispatched largest first;
//    the hardware dispatcher hands the remaining runs to CUs as they free up.
//  * otherwise: a cooperative launch that fills the GPU; every run gets at
//    least one workgroup and the remaining ones go to the runs with the most
//    work per workgroup (cells^1.5 ~ cells x steps), so the largest meshes no
//    longer set the pace alone.  A cluster splits its run's edges and cells
//    into contiguous (Hilbert-ordered) ranges and meets at three barriers per
//    step (after the fluxes, the CFL partials and the update): agent-scope
//    release / acquire around a per-run arrival counter.  Clusters are packed
//    XCD by XCD (workgroup w runs on XCD w mod 8), so a run's workgroups
//    share one L2.  A barrier wait has a spin limit; reaching it raises an
//    abort flag that every workgroup polls, so the launch always drains.
//
// Layout: all runs' arrays concatenated (cells, internal edges, boundary
// edges, incidences, series), each run's in the device order of
// build_tables (Hilbert cells, edges by lower cell), indices local to the run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"
#include "sp_adjoint.h"
#include "sp_physics.h"
#include "sw2dg_internal.h"

using namespace sw2dg_detail;
using sp::kSmall;

namespace {

constexpr int kMaxSeries = 8;

struct RunDesc {
    int nc, ni, nb;      // sizes
    int c0, e0, b0;      // offsets of cells, internal edges, boundary edges
    int ip0;             // offset of the run's int_ptr / bnd_ptr (nc + 1 entries each)
    int ns, s0;          // series count and offset (types)
    int nt, t0;          // knots count and offset (times); values at s0 * nt_max? see v0
    int v0;              // offset of the run's series values [ns][nt]
    int ng, g0;          // gauges
};

struct EnsDev {
    const RunDesc *runs;
    // cells
    const double *zb, *area, *phi, *beta, *K1, *K2, *fcos, *fsin;
    double *h, *qx, *qy;
    const int *int_ptr, *int_inc, *bnd_ptr, *bnd_inc;
    // internal edges
    const int *eL, *eR;
    const double *elen, *enx, *eny;
    double *F, *sw;  // F: 5 per internal edge, then 3 per boundary edge (offset Fb)
    size_t Fb;
    // boundary edges
    const int *bcell, *bser;
    const double *blen, *bnx, *bny;
    // series
    const int *stype;
    const double *st, *sv;
    // gauges
    const int *gcell;
};

__device__ double interp_e_(const double *tt, const double *vv, int n, double t) {
    if (n == 0) return 0.0;
    if (n == 1 || t <= tt[0] - 1.0e-12) return vv[0];
    if (t >= tt[n - 1]) return vv[n - 1];
    int it = 0;
    while (it < n - 1 && t > tt[it + 1]) it++;
    const double t1 = tt[it], t2 = tt[it + 1], v1 = vv[it], v2 = vv[it + 1];
    const double den = t2 - t1;
    if (den < 1e-12) return v1;
    return v1 + (v2 - v1) * (t - t1) / den;
}

// workgroup -> (run, rank in the run's cluster); per-run cluster size,
// arrival counter and CFL partial minima
struct Clusters {
    const int *wg_run, *wg_rank, *run_c;
    unsigned *arrive;
    double *part;  // [run][cmax]
    int cmax;
    int *abort;
};

constexpr unsigned kSpinLimit = 1u << 24;  // ~1 s of s_sleep 2 per wait

// barrier of the c workgroups of one run (c == 1: the workgroup barrier)
__device__ __forceinline__ bool run_sync(const Clusters &cl, int r, int c, unsigned &target, int *s_flag) {
    __syncthreads();
    if (c == 1) return true;
    if (threadIdx.x == 0) {
        target += (unsigned)c;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(cl.arrive + r, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(cl.arrive + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (__hip_atomic_load(cl.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (++spins > kSpinLimit) {
                __hip_atomic_store(cl.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0;
}

template <int B>
__global__ __launch_bounds__(B) void k_ens(EnsDev d, Clusters cl, Opts o, double t_start,
                                           double t_max, double dt_max, double cfl, int n_obs,
                                           const double *__restrict__ obs_t, int ng_max,
                                           double *__restrict__ gser, int64_t max_steps,
                                           int64_t *steps_out) {
    const int r = cl.wg_run[blockIdx.x];
    if (r < 0) return;  // unused slot of the XCD packing
    const int rank = cl.wg_rank[blockIdx.x], c = cl.run_c[r];
    const RunDesc R = d.runs[r];
    const int tid = (int)threadIdx.x;
    const int nc = R.nc, ni = R.ni, nb = R.nb;
    // this workgroup's contiguous share of the run's edges and cells
    const int e_lo = (int)((int64_t)(ni + nb) * rank / c), e_hi = (int)((int64_t)(ni + nb) * (rank + 1) / c);
    const int c_lo = (int)((int64_t)nc * rank / c), c_hi = (int)((int64_t)nc * (rank + 1) / c);
    double *part = cl.part + (size_t)r * cl.cmax;
    // run-local views
    const double *zb = d.zb + R.c0, *area = d.area + R.c0, *phi = d.phi + R.c0;
    const double *beta = d.beta + R.c0, *K1 = d.K1 + R.c0, *K2 = d.K2 + R.c0;
    const double *fcos = d.fcos + R.c0, *fsin = d.fsin + R.c0;
    double *h = d.h + R.c0, *qx = d.qx + R.c0, *qy = d.qy + R.c0;
    const int *iptr = d.int_ptr + R.ip0, *bptr = d.bnd_ptr + R.ip0;
    const int *iinc = d.int_inc + 2 * (size_t)R.e0, *binc = d.bnd_inc + R.b0;
    const int *eL = d.eL + R.e0, *eR = d.eR + R.e0;
    const double *elen = d.elen + R.e0, *enx = d.enx + R.e0, *eny = d.eny + R.e0;
    double *F = d.F + 5 * (size_t)R.e0, *Fb = d.F + d.Fb + 3 * (size_t)R.b0;
    double *sw = d.sw + 2 * (size_t)R.e0;
    const int *bcell = d.bcell + R.b0, *bser = d.bser + R.b0;
    const double *blen = d.blen + R.b0, *bnx = d.bnx + R.b0, *bny = d.bny + R.b0;
    const int *stype = d.stype + R.s0;
    const double *st = d.st + R.t0, *sv = d.sv + R.v0;
    const int *gcell = d.gcell + R.g0;
    double *gout = gser + (size_t)r * n_obs * ng_max;

    __shared__ double red[B / 64];
    __shared__ double s_bcv[kMaxSeries];
    __shared__ double s_dt;
    __shared__ int s_trig, s_oi, s_flag;
    double t = t_start;
    int oi = 0;
    int64_t step = 0;
    unsigned target = 0;
    while (t < t_max && step < max_steps) {
        if (tid < R.ns) s_bcv[tid] = interp_e_(st, sv + (size_t)tid * R.nt, R.nt, t);
        __syncthreads();
        // ---- edge fluxes (as the single-run dry run's)
        for (int e = e_lo + tid; e < e_hi; e += B) {
            if (e >= ni) {
                const int j = e - ni;
                const int i = bcell[j];
                const spe::Side sd = spe::lr_fwd(h[i], qx[i], qy[i], zb[i], beta[i], bnx[j], bny[j], o.h_min);
                const int ser = bser[j];
                double f[3];
                spe::bnd_fwd(o.h_min, o.phi_min, beta[i], phi[i], zb[i], sd.z, sd.q, sd.r, sd.lp,
                             stype[ser], s_bcv[ser], f);
                double *fb = Fb + 3 * (size_t)j;
                fb[0] = f[0];
                fb[1] = f[1];
                fb[2] = f[2];
                continue;
            }
            const int L = eL[e], Rr = eR[e];
            const double nx = enx[e], ny = eny[e];
            const double hL = h[L], hR = h[Rr];
            const spe::Side sL = spe::lr_fwd(hL, qx[L], qy[L], zb[L], beta[L], nx, ny, o.h_min);
            const spe::Side sR = spe::lr_fwd(hR, qx[Rr], qy[Rr], zb[Rr], beta[Rr], nx, ny, o.h_min);
            double f[5];
            spe::hllc_fwd(o.h_min, o.phi_min, beta[L], beta[Rr], phi[L], phi[Rr], zb[L], zb[Rr], sL.z, sL.q,
                          sL.r, sR.z, sR.q, sR.r, sL.lm, sR.lm, sL.lp, sR.lp, f);
            // F1 and each side's momentum sum with its cell's factor (the
            // sums of side_contrib; the update multiplies by dt)
            const double len = elen[e];
            const double coL = spe::edge_coeff(len, area[L], phi[L], o.phi_min);
            const double coR = spe::edge_coeff(len, area[Rr], phi[Rr], o.phi_min);
            double *Fe = F + 5 * (size_t)e;
            Fe[0] = f[0];
            Fe[1] = (-coL * nx) * f[1] + (coL * nx) * f[3] + (coL * ny) * f[2];
            Fe[2] = (-coL * ny) * f[1] + (coL * ny) * f[3] + (-coL * nx) * f[2];
            Fe[3] = (coR * nx) * f[1] + (coR * nx) * f[4] + (-coR * ny) * f[2];
            Fe[4] = (coR * ny) * f[1] + (coR * ny) * f[4] + (coR * nx) * f[2];
            double m1 = fabs(sL.lm), v = fabs(sL.lp);
            if (m1 < v) m1 = v;
            if (m1 < (v = fabs(sR.lm))) m1 = v;
            if (m1 < (v = fabs(sR.lp))) m1 = v;
            reinterpret_cast<double2 *>(sw)[e] =
                make_double2((hL > o.h_min) ? m1 * len : 0.0, (hR > o.h_min) ? m1 * len : 0.0);
        }
        if (!run_sync(cl, r, c, target, &s_flag)) break;
        // ---- CFL (k_e_cfl): per-cell candidate, cluster minimum
        double cand = dt_max;
        for (int i = c_lo + tid; i < c_hi; i += B) {
            const double hi = h[i];
            double swept = 0.0;
            if (hi > o.h_min) {
                for (int k = bptr[i]; k < bptr[i + 1]; ++k) {
                    const int j = binc[k];
                    const spe::Side s = spe::lr_fwd(hi, qx[i], qy[i], zb[i], beta[i], bnx[j], bny[j], o.h_min);
                    const double lm = (fabs(s.lp) < fabs(s.lm)) ? fabs(s.lm) : fabs(s.lp);
                    swept += lm * blen[j];
                }
            }
            const int ie = iptr[i + 1];
            for (int kb = iptr[i]; kb < ie; kb += 4) {  // four incidences' loads in flight, summed in order
                int in[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) in[j] = (kb + j < ie) ? iinc[kb + j] : 0;
                double w[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) w[j] = (kb + j < ie) ? sw[in[j]] : 0.0;  // (edge << 1) | side
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (kb + j < ie) swept += w[j];
            }
            const double den = (swept < kSmall) ? kSmall : swept;
            const double x = cfl * (area[i] / den);
            if (x < cand) cand = x;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o2 = __shfl_xor(cand, off, 64);
            cand = (o2 < cand) ? o2 : cand;
        }
        if ((tid & 63) == 0) red[tid >> 6] = cand;
        __syncthreads();
        if (tid == 0) {
            double m = red[0];
            for (int w = 1; w < B / 64; ++w) m = (red[w] < m) ? red[w] : m;
            part[rank] = m;
        }
        if (!run_sync(cl, r, c, target, &s_flag)) break;
        // ---- step schedule (k_e_dt), replicated in every workgroup of the run
        if (tid == 0) {
            double dtc = part[0];
            for (int w = 1; w < c; ++w) dtc = (part[w] < dtc) ? part[w] : dtc;
            if ((t + dtc) > t_max) dtc = t_max - t;
            if (dtc < 1e-6) dtc = 1e-6;
            double target = dt_max;
            int oo = oi;
            if (oo < n_obs) {
                const double to = obs_t[oo] - t;
                if (to - 10e-6 > 0.0 && to < target) target = to;
            }
            double dt = (dtc < target) ? dtc : target;
            if ((t + dt) > t_max) dt = t_max - t;
            if (dt < 1e-6) dt = 1e-6;
            const double tn = t + dt;
            while (oo < n_obs && tn > obs_t[oo] + 1e-6) oo++;
            int tr = -1;
            if (oo < n_obs && fabs(tn - obs_t[oo]) < 1e-6) tr = oo++;
            s_dt = dt;
            s_trig = tr;
            s_oi = oo;
        }
        __syncthreads();
        const double dt = s_dt;
        // ---- cell update (k_e_update): balance, clean-up, clamp, friction
        for (int i = c_lo + tid; i < c_hi; i += B) {
            double sh = h[i], sx = qx[i], sy = qy[i];
            const double ph = phi[i], ar = area[i];
            for (int k = bptr[i]; k < bptr[i + 1]; ++k) {
                const int j = binc[k];
                const double nx = bnx[j], ny = bny[j];
                const double *f = Fb + 3 * (size_t)j;
                const double co = spe::edge_coeff(blen[j], ar, ph, o.phi_min);
                sh += (+co) * f[0] * dt;
                sx += ((+co * nx) * f[1] + (+co * nx) * 0.0 + (-co * ny) * f[2]) * dt;
                sy += ((+co * ny) * f[1] + (+co * ny) * 0.0 + (+co * nx) * f[2]) * dt;
            }
            const int ie = iptr[i + 1];
            for (int kb = iptr[i]; kb < ie; kb += 2) {  // two incidences' loads together, sums in order
                int in[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) in[j] = (kb + j < ie) ? iinc[kb + j] : 0;
                double len[2], f0[2], f1[2], f2[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (kb + j < ie) {
                        const int e = in[j] >> 1, side = in[j] & 1;
                        const double *f = F + 5 * (size_t)e;
                        len[j] = elen[e];
                        f0[j] = f[0];
                        f1[j] = f[1 + 2 * side];
                        f2[j] = f[2 + 2 * side];
                    }
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (kb + j < ie) {
                        const double co = spe::edge_coeff(len[j], ar, ph, o.phi_min);
                        sh += ((in[j] & 1) ? co : -co) * (f0[j] * dt);
                        sx += f1[j] * dt;
                        sy += f2[j] * dt;
                    }
                }
            }
            if (fabs(sh) < kSmall) sh = 0.0;
            if (fabs(sx) < kSmall) sx = 0.0;
            if (fabs(sy) < kSmall) sy = 0.0;
            if (sh < 0.) {
                sh = 0.;
                sx = 0.;
                sy = 0.;
            }
            if (o.friction != sp::FRIC_NONE)
                spe::fric_fwd(o.h_min, sh, sx, sy, fcos[i], fsin[i], K1[i], K2[i], dt);
            h[i] = sh;
            qx[i] = sx;
            qy[i] = sy;
        }
        const int tr = s_trig;
        t = t + dt;
        oi = s_oi;
        ++step;
        if (!run_sync(cl, r, c, target, &s_flag)) break;
        // ---- gauges of an observation step (the cluster's first workgroup)
        if (tr >= 0 && rank == 0)
            for (int g = tid; g < R.ng; g += B) gout[(size_t)tr * ng_max + g] = h[gcell[g]];
    }
    if (tid == 0 && rank == 0) steps_out[r] = step;
}

}  // namespace

struct sw2dg_ens {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    sw2dg_params prm{};
    Opts o{};
    int n_runs = 0;
    std::vector<MeshTables> T;
    std::vector<RunDesc> runs;
    int64_t NC = 0, NI = 0, NB = 0, NIP = 0;
    std::vector<char> run_set;
    // host staging of the per-run inputs (device order); series and gauges
    // per run, concatenated at launch
    std::vector<double> phi, beta, K1, K2, fcos, fsin, h0, qx0, qy0;
    std::vector<std::vector<int>> r_type, r_gauge;
    std::vector<std::vector<double>> r_t, r_v;
    EnsDev d{};
    int max_cluster = 64;  // workgroups per run in the cooperative mode (1: one per run)
    int last_wgs = 0, last_cmax = 1;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    template <class T_>
    T_ *dalloc(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T_)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        bytes += (int64_t)(std::max<size_t>(n, 1) * sizeof(T_));
        return (T_ *)p;
    }
    template <class T_>
    T_ *upload(const std::vector<T_> &v) {
        T_ *p = dalloc<T_>(v.size());
        if (p && !v.empty()) hipMemcpy(p, v.data(), v.size() * sizeof(T_), hipMemcpyHostToDevice);
        return p;
    }
    ~sw2dg_ens() {
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
int efail(sw2dg_ens *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}
#define EHIP(e, x)                                                                      \
    do {                                                                                \
        hipError_t err_ = (x);                                                          \
        if (err_ != hipSuccess) return efail(e, SW2DG_ERR_HIP, hipGetErrorString(err_)); \
    } while (0)

// Cluster plan: workgroups per run and the workgroup -> (run, rank) tables.
// capacity = co-resident workgroups of the launch (0: one workgroup per run,
// largest first).  In the cooperative mode every run gets one workgroup, the
// rest go one by one to the run with the largest work per workgroup
// (cells^1.5), at most max_c per run and at least 2 B cells per workgroup;
// clusters are then packed into the 8 XCDs (first fit, largest first;
// workgroup w sits on XCD w mod 8) and unused slots get run -1.
struct Plan {
    std::vector<int> wg_run, wg_rank, run_c;
    int cmax = 1;
};

Plan make_plan(const std::vector<RunDesc> &runs, int capacity, int max_c, int B) {
    const int n = (int)runs.size();
    Plan p;
    p.run_c.assign(n, 1);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return runs[a].nc + runs[a].ni > runs[b].nc + runs[b].ni; });
    if (capacity <= 0 || n > capacity) {
        p.wg_run = order;
        p.wg_rank.assign(n, 0);
        return p;
    }
    std::vector<double> w(n);
    std::vector<int> cap(n);
    for (int r = 0; r < n; ++r) {
        w[r] = std::pow((double)runs[r].nc, 1.5);
        cap[r] = std::max(1, std::min(max_c, runs[r].nc / (2 * B)));
    }
    std::priority_queue<std::pair<double, int>> q;
    for (int r = 0; r < n; ++r)
        if (cap[r] > 1) q.push({w[r], r});
    int used = n;
    while (used < capacity && !q.empty()) {
        const int r = q.top().second;
        q.pop();
        p.run_c[r] += 1;
        ++used;
        if (p.run_c[r] < cap[r]) q.push({w[r] / p.run_c[r], r});
    }
    // XCD packing
    const int nx = 8, per = capacity / nx;
    std::vector<std::vector<std::pair<int, int>>> bins(nx);  // (run, rank)
    std::vector<int> order_c(n);
    std::iota(order_c.begin(), order_c.end(), 0);
    std::stable_sort(order_c.begin(), order_c.end(), [&](int a, int b) { return p.run_c[a] > p.run_c[b]; });
    for (int r : order_c) {
        int c = p.run_c[r], k = 0;
        int fit = -1;
        for (int x = 0; x < nx; ++x)
            if ((int)bins[x].size() + c <= per) {
                fit = x;
                break;
            }
        while (k < c) {
            int x = fit;
            if (x < 0) {  // split over the emptiest bins
                x = 0;
                for (int y = 1; y < nx; ++y)
                    if (bins[y].size() < bins[x].size()) x = y;
            }
            bins[x].push_back({r, k++});
            if (fit >= 0 && (int)bins[fit].size() >= per) fit = -1;
        }
        p.cmax = std::max(p.cmax, c);
    }
    size_t longest = 0;
    for (auto &b : bins) longest = std::max(longest, b.size());
    const int G = (int)(longest * nx);
    p.wg_run.assign(G, -1);
    p.wg_rank.assign(G, 0);
    for (int x = 0; x < nx; ++x)
        for (size_t s_ = 0; s_ < bins[x].size(); ++s_) {
            p.wg_run[s_ * nx + x] = bins[x][s_].first;
            p.wg_rank[s_ * nx + x] = bins[x][s_].second;
        }
    return p;
}

template <int B>
int co_resident(sw2dg_ens *E) {
    int per_cu = 0, n_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ens<B>, B, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, E->device) != hipSuccess) return 0;
    return per_cu * n_cu;
}
}  // namespace

extern "C" {

int sw2dg_ens_create(int32_t n_runs, const sw2dg_mesh *meshes, const sw2dg_params *prm,
                     int32_t device, sw2dg_ens **out) {
    if (!out) return SW2DG_ERR_ARG;
    *out = nullptr;
    if (n_runs <= 0 || !meshes || !prm) return SW2DG_ERR_ARG;
    if (prm->scheme != SW2DG_GODUNOV || prm->head_loss) return SW2DG_ERR_ARG;
    int64_t NC = 0, NI = 0, NB = 0;
    for (int r = 0; r < n_runs; ++r) {
        const sw2dg_mesh *m = meshes + r;
        if (m->n_cell <= 0 || m->n_int < 0 || m->n_bnd < 0) return SW2DG_ERR_ARG;
        for (int64_t e = 0; e < m->n_bnd; ++e)
            if (m->bnd_cell[e] < 0 || m->bnd_cell[e] >= m->n_cell || m->bnd_series[e] < 0 ||
                m->bnd_series[e] >= kMaxSeries)
                return SW2DG_ERR_ARG;
        for (int64_t e = 0; e < m->n_int; ++e)
            if (m->int_L[e] < 0 || m->int_L[e] >= m->n_cell || m->int_R[e] < 0 ||
                m->int_R[e] >= m->n_cell || m->int_L[e] == m->int_R[e])
                return SW2DG_ERR_ARG;
        NC += m->n_cell;
        NI += m->n_int;
        NB += m->n_bnd;
    }
    if (NC > INT32_MAX / 2 || NI > INT32_MAX / 8) return SW2DG_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SW2DG_ERR_NO_DEVICE;
    auto *E = new sw2dg_ens();
    if (device < 0) hipGetDevice(&E->device);
    else E->device = device;
    if (hipSetDevice(E->device) != hipSuccess || hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
        delete E;
        return SW2DG_ERR_HIP;
    }
    E->prm = *prm;
    E->o.scheme = prm->scheme;
    E->o.topo = prm->topo;
    E->o.friction = prm->friction;
    E->o.head_loss = 0;
    E->o.h_min = prm->h_min;
    E->o.phi_min = prm->phi_min;
    E->n_runs = n_runs;
    E->NC = NC;
    E->NI = NI;
    E->NB = NB;
    E->T.resize(n_runs);
    E->runs.resize(n_runs);
    E->run_set.assign(n_runs, 0);
    E->r_type.resize(n_runs);
    E->r_gauge.resize(n_runs);
    E->r_t.resize(n_runs);
    E->r_v.resize(n_runs);
    std::vector<double> zb(NC), area(NC), elen(NI), enx(NI), eny(NI), blen(NB), bnx(NB), bny(NB);
    std::vector<int> iptr, bptr, iinc(2 * NI), binc(NB), eL(NI), eR(NI), bcell(NB), bser(NB);
    int64_t c0 = 0, e0 = 0, b0 = 0;
    for (int r = 0; r < n_runs; ++r) {
        const sw2dg_mesh *m = meshes + r;
        MeshTables &T = E->T[r];
        build_tables(m, T);
        RunDesc &R = E->runs[r];
        R.nc = (int)m->n_cell;
        R.ni = (int)m->n_int;
        R.nb = (int)m->n_bnd;
        R.c0 = (int)c0;
        R.e0 = (int)e0;
        R.b0 = (int)b0;
        R.ip0 = (int)iptr.size();
        for (int64_t k = 0; k < m->n_cell; ++k) {
            zb[c0 + k] = m->cell_zb[T.perm[k]];
            area[c0 + k] = m->cell_area[T.perm[k]];
        }
        iptr.insert(iptr.end(), T.iptr.begin(), T.iptr.end());
        bptr.insert(bptr.end(), T.bptr.begin(), T.bptr.end());
        std::copy(T.iinc.begin(), T.iinc.end(), iinc.begin() + 2 * e0);
        std::copy(T.binc.begin(), T.binc.end(), binc.begin() + b0);
        std::copy(T.eL.begin(), T.eL.end(), eL.begin() + e0);
        std::copy(T.eR.begin(), T.eR.end(), eR.begin() + e0);
        std::copy(T.elen.begin(), T.elen.end(), elen.begin() + e0);
        std::copy(T.enx.begin(), T.enx.end(), enx.begin() + e0);
        std::copy(T.eny.begin(), T.eny.end(), eny.begin() + e0);
        std::copy(T.blen.begin(), T.blen.end(), blen.begin() + b0);
        std::copy(T.bnx.begin(), T.bnx.end(), bnx.begin() + b0);
        std::copy(T.bny.begin(), T.bny.end(), bny.begin() + b0);
        std::copy(T.bser.begin(), T.bser.begin() + m->n_bnd, bser.begin() + b0);
        std::copy(T.bcell.begin(), T.bcell.begin() + m->n_bnd, bcell.begin() + b0);
        c0 += m->n_cell;
        e0 += m->n_int;
        b0 += m->n_bnd;
    }
    E->NIP = (int64_t)iptr.size();
    E->phi.assign(NC, 1.0);
    E->beta.assign(NC, 1.0);
    E->K1.assign(NC, 0.0);
    E->K2.assign(NC, 0.0);
    E->fcos.assign(NC, 1.0);
    E->fsin.assign(NC, 0.0);
    E->h0.assign(NC, 0.0);
    E->qx0.assign(NC, 0.0);
    E->qy0.assign(NC, 0.0);
    EnsDev &d = E->d;
    d.zb = E->upload(zb);
    d.area = E->upload(area);
    d.int_ptr = E->upload(iptr);
    d.bnd_ptr = E->upload(bptr);
    d.int_inc = E->upload(iinc);
    d.bnd_inc = E->upload(binc);
    d.eL = E->upload(eL);
    d.eR = E->upload(eR);
    d.elen = E->upload(elen);
    d.enx = E->upload(enx);
    d.eny = E->upload(eny);
    d.blen = E->upload(blen);
    d.bnx = E->upload(bnx);
    d.bny = E->upload(bny);
    d.bser = E->upload(bser);
    d.bcell = E->upload(bcell);
    d.h = E->dalloc<double>(NC);
    d.qx = E->dalloc<double>(NC);
    d.qy = E->dalloc<double>(NC);
    d.Fb = 5 * (size_t)std::max<int64_t>(NI, 1);
    d.F = E->dalloc<double>(d.Fb + 3 * (size_t)std::max<int64_t>(NB, 1));
    d.sw = E->dalloc<double>(2 * (size_t)std::max<int64_t>(NI, 1));
    for (void *p : E->allocs)
        if (!p) {
            delete E;
            return SW2DG_ERR_HIP;
        }
    *out = E;
    return SW2DG_OK;
}

int sw2dg_ens_set_run(sw2dg_ens *E, int32_t run, const double *phi, const double *K1,
                      const double *K2, const double *fric_alpha_deg, int32_t n_series,
                      int32_t n_times, const int32_t *type, const double *t, const double *v,
                      const double *h, const double *qx, const double *qy, int32_t n_gauges,
                      const int64_t *gauge_cell) {
    if (!E || run < 0 || run >= E->n_runs) return SW2DG_ERR_ARG;
    if (!phi || !K1 || !K2 || !h || !qx || !qy) return efail(E, SW2DG_ERR_ARG, "missing run input");
    if (n_series < 0 || n_series > kMaxSeries || (n_series > 0 && (n_times < 1 || !type || !t || !v)))
        return efail(E, SW2DG_ERR_ARG, "bad boundary series");
    if (n_gauges < 0 || (n_gauges > 0 && !gauge_cell)) return efail(E, SW2DG_ERR_ARG, "bad gauges");
    const MeshTables &T = E->T[run];
    RunDesc &R = E->runs[run];
    for (int64_t e = 0; e < T.nb; ++e)
        if (T.bser[e] >= n_series) return efail(E, SW2DG_ERR_ARG, "a boundary edge refers to a missing series");
    for (int k = 0; k < n_series; ++k)
        if (type[k] < 0 || type[k] > 3) return efail(E, SW2DG_ERR_BC, "BC type not handled");
    for (int g = 0; g < n_gauges; ++g)
        if (gauge_cell[g] < 0 || gauge_cell[g] >= T.nc) return efail(E, SW2DG_ERR_ARG, "gauge cell out of range");
    const double rad = std::atan(1.0) / 45.0;
    for (int64_t k = 0; k < T.nc; ++k) {
        const int src = T.perm[k];
        const size_t q = R.c0 + k;
        E->phi[q] = phi[src];
        E->K1[q] = K1[src];
        E->K2[q] = K2[src];
        const double a = fric_alpha_deg ? fric_alpha_deg[src] * rad : 0.0;
        E->fcos[q] = std::cos(a);
        E->fsin[q] = std::sin(a);
        E->h0[q] = h[src];
        E->qx0[q] = qx[src];
        E->qy0[q] = qy[src];
    }
    R.ns = n_series;
    R.nt = n_series > 0 ? n_times : 0;
    R.ng = n_gauges;
    E->r_type[run].assign(type, type + n_series);
    E->r_t[run].assign(t, t + R.nt);
    E->r_v[run].assign(v, v + (size_t)n_series * R.nt);
    E->r_gauge[run].resize(n_gauges);
    for (int g = 0; g < n_gauges; ++g) E->r_gauge[run][g] = T.inv[gauge_cell[g]];
    E->run_set[run] = 1;
    return SW2DG_OK;
}

int sw2dg_ens_run(sw2dg_ens *E, int32_t n_obs, const double *obs_t, int32_t block,
                  int64_t max_steps, double *gauge_out, int64_t *steps_out, double *device_ms) {
    if (!E || n_obs < 0 || (n_obs > 0 && !obs_t) || max_steps <= 0) return SW2DG_ERR_ARG;
    if (block != 256 && block != 512 && block != 1024) return efail(E, SW2DG_ERR_ARG, "block must be 256, 512 or 1024");
    for (int r = 0; r < E->n_runs; ++r)
        if (!E->run_set[r]) return efail(E, SW2DG_ERR_ARG, "sw2dg_ens_set_run not called for every run");
    EHIP(E, hipSetDevice(E->device));
    EnsDev &d = E->d;
    // concatenate the runs' series and gauges
    std::vector<int> c_type, c_gauge;
    std::vector<double> c_t, c_v;
    int ng_max = 0;
    for (int r = 0; r < E->n_runs; ++r) {
        RunDesc &R = E->runs[r];
        R.s0 = (int)c_type.size();
        c_type.insert(c_type.end(), E->r_type[r].begin(), E->r_type[r].end());
        R.t0 = (int)c_t.size();
        c_t.insert(c_t.end(), E->r_t[r].begin(), E->r_t[r].end());
        R.v0 = (int)c_v.size();
        c_v.insert(c_v.end(), E->r_v[r].begin(), E->r_v[r].end());
        R.g0 = (int)c_gauge.size();
        c_gauge.insert(c_gauge.end(), E->r_gauge[r].begin(), E->r_gauge[r].end());
        ng_max = std::max(ng_max, R.ng);
    }
    std::vector<void *> tmp;
    auto up = [&](auto &v) {
        using T_ = typename std::decay_t<decltype(v)>::value_type;
        T_ *p = nullptr;
        if (hipMalloc((void **)&p, std::max<size_t>(v.size(), 1) * sizeof(T_)) != hipSuccess) return (T_ *)nullptr;
        tmp.push_back(p);
        if (!v.empty()) hipMemcpy(p, v.data(), v.size() * sizeof(T_), hipMemcpyHostToDevice);
        return p;
    };
    d.phi = up(E->phi);
    d.beta = up(E->beta);
    d.K1 = up(E->K1);
    d.K2 = up(E->K2);
    d.fcos = up(E->fcos);
    d.fsin = up(E->fsin);
    d.stype = up(c_type);
    d.st = up(c_t);
    d.sv = up(c_v);
    d.gcell = up(c_gauge);
    d.runs = up(E->runs);
    std::vector<double> ot(obs_t, obs_t + n_obs);
    double *d_obs = up(ot);
    int capacity = 0;
    if (E->max_cluster > 1)
        capacity = block == 256 ? co_resident<256>(E) : block == 512 ? co_resident<512>(E) : co_resident<1024>(E);
    const Plan plan = make_plan(E->runs, capacity, E->max_cluster, block);
    const bool coop = capacity > 0 && E->n_runs <= capacity;
    E->last_wgs = (int)plan.wg_run.size();
    E->last_cmax = plan.cmax;
    Clusters cl{};
    cl.wg_run = up(plan.wg_run);
    cl.wg_rank = up(plan.wg_rank);
    cl.run_c = up(plan.run_c);
    cl.cmax = plan.cmax;
    std::vector<unsigned> zeros_u(E->n_runs, 0u);
    cl.arrive = up(zeros_u);
    std::vector<double> part((size_t)E->n_runs * plan.cmax, 0.0);
    cl.part = up(part);
    std::vector<int> abort_v(1, 0);
    cl.abort = up(abort_v);
    const size_t ng = (size_t)E->n_runs * std::max(1, n_obs) * std::max(1, ng_max);
    double *d_g = nullptr;
    int64_t *d_steps = nullptr;
    bool ok = hipMalloc((void **)&d_g, ng * sizeof(double)) == hipSuccess;
    if (ok) tmp.push_back(d_g);
    ok = ok && hipMalloc((void **)&d_steps, E->n_runs * sizeof(int64_t)) == hipSuccess;
    if (ok) tmp.push_back(d_steps);
    for (void *p : tmp)
        if (!p) ok = false;
    int st = SW2DG_OK;
    if (!ok) st = efail(E, SW2DG_ERR_HIP, "out of device memory for the ensemble inputs");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st) {
        hipMemsetAsync(d_g, 0, ng * sizeof(double), E->stream);
        hipMemcpyAsync(d.h, E->h0.data(), E->NC * sizeof(double), hipMemcpyHostToDevice, E->stream);
        hipMemcpyAsync(d.qx, E->qx0.data(), E->NC * sizeof(double), hipMemcpyHostToDevice, E->stream);
        hipMemcpyAsync(d.qy, E->qy0.data(), E->NC * sizeof(double), hipMemcpyHostToDevice, E->stream);
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipEventRecord(e0, E->stream);
        const sw2dg_params &p = E->prm;
        const dim3 grid((unsigned)plan.wg_run.size());
        double t0 = p.t0, tm = p.t_max, dtm = p.dt_max, cf = p.cfl_max;
        int ngm = std::max(1, ng_max);
        int64_t ms_ = max_steps;
        Opts o = E->o;
        void *args[] = {&d, &cl, &o, &t0, &tm, &dtm, &cf, &n_obs, &d_obs, &ngm, &d_g, &ms_, &d_steps};
        const void *fn = block == 256 ? (const void *)k_ens<256> : block == 512 ? (const void *)k_ens<512>
                                                                                 : (const void *)k_ens<1024>;
        hipError_t err = coop ? hipLaunchCooperativeKernel(fn, grid, dim3(block), args, 0, E->stream)
                              : hipLaunchKernel(fn, grid, dim3(block), args, 0, E->stream);
        hipEventRecord(e1, E->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(E->stream);
        if (err == hipSuccess) err = hipGetLastError();
        if (err != hipSuccess) st = efail(E, SW2DG_ERR_HIP, hipGetErrorString(err));
        int aborted = 0;
        if (!st && hipMemcpy(&aborted, cl.abort, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            st = efail(E, SW2DG_ERR_HIP, "copy of the abort flag failed");
        if (!st && aborted) st = efail(E, SW2DG_ERR_HIP, "a run's workgroups did not meet at a barrier (spin limit)");
        float ms = 0.f;
        if (!st) hipEventElapsedTime(&ms, e0, e1);
        if (device_ms) *device_ms = ms;
        if (!st && gauge_out)
            if (hipMemcpy(gauge_out, d_g, ng * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                st = efail(E, SW2DG_ERR_HIP, "copy of the gauge series failed");
        if (!st && steps_out)
            if (hipMemcpy(steps_out, d_steps, E->n_runs * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
                st = efail(E, SW2DG_ERR_HIP, "copy of the step counts failed");
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    for (void *p : tmp) (void)hipFree(p);
    return st;
}

int sw2dg_ens_get_state(sw2dg_ens *E, int32_t run, double *h, double *qx, double *qy) {
    if (!E || run < 0 || run >= E->n_runs) return SW2DG_ERR_ARG;
    EHIP(E, hipSetDevice(E->device));
    const RunDesc &R = E->runs[run];
    const MeshTables &T = E->T[run];
    std::vector<double> v(R.nc);
    double *dst[3] = {h, qx, qy};
    double *src[3] = {E->d.h, E->d.qx, E->d.qy};
    for (int f = 0; f < 3; ++f) {
        if (!dst[f]) continue;
        EHIP(E, hipMemcpy(v.data(), src[f] + R.c0, R.nc * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < R.nc; ++k) dst[f][T.perm[k]] = v[k];
    }
    return SW2DG_OK;
}

int sw2dg_ens_set_cluster_limit(sw2dg_ens *E, int32_t max_wg_per_run) {
    if (!E || max_wg_per_run < 1 || max_wg_per_run > 1024) return SW2DG_ERR_ARG;
    E->max_cluster = max_wg_per_run;
    return SW2DG_OK;
}

int sw2dg_ens_last_launch(sw2dg_ens *E, int32_t *workgroups, int32_t *max_wg_per_run) {
    if (!E) return SW2DG_ERR_ARG;
    if (workgroups) *workgroups = E->last_wgs;
    if (max_wg_per_run) *max_wg_per_run = E->last_cmax;
    return SW2DG_OK;
}

int64_t sw2dg_ens_device_bytes(sw2dg_ens *E) { return E ? E->bytes : 0; }
const char *sw2dg_ens_last_error(sw2dg_ens *E) { return E ? E->err.c_str() : "null ensemble"; }
void sw2dg_ens_destroy(sw2dg_ens *E) { delete E; }

}  // extern "C"
