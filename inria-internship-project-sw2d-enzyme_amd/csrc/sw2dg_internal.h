// sw2dg_internal.h -- device tables and the context shared by the forward
// (sw2dg_forward.hip) and adjoint (sw2dg_adjoint.hip) translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"

namespace sw2dg_detail {

constexpr int kBlock = 256;
constexpr int kDtBlock = 1024;  // the one-block dt reductions

// device-resident clock of the native forward (one per ctx).  Slot p holds the
// values valid at the start of a step launched with parity p; the step
// writes slot 1-p.
struct Clock {
    double T[2];         // time at the start of the step
    double DT[2];        // dt of the previous step (MUSCL reconstruction)
    long long STEP[2];   // steps taken since set_state
    int ACT[2];          // step runs (1) or is a no-op (0)
    int err;             // first error code raised on the device
    int store_flags;     // storage instants hit by the last step
    int paused;          // a storage instant stopped the run
    int pause_on_store;
    double t_stop, t_max, dt_max, dt_map, dt_ts, cfl;
    long long step_stop;
    double dt_last;      // dt of the last active step
    int n_series;
    double t_half;       // source interpolation time of the step (t + dt/2)
    int wind_k;          // record bracket at t_half: -1 first record, -2 last, k >= 1 (k-1, k)
    int rain_k;          // the same on the rain records
    double parab_dt;     // explicit momentum-diffusion substep of the step
    int parab_n;         // its number of substeps
    int parab_pending;   // n > 1: substeps left to the host (the launch's later steps idle)
    int act_saved;       // ACT of the next step, restored once the substeps are done
};

struct BcDev {
    const double *t;     // [n_times]
    const double *v;     // [n_series*n_times]
    const int *type;     // [n_series]
    double *vinf, *vsup, *tinf, *tsup, *val;
    int *rank, *atend;
    int n_times;
    int sorted;          // t is non-decreasing: the _E interpolation may bisect
};

struct MeshDev {
    int n_cell, n_bnd, n_int;
    // cells (renumbered)
    double *h, *qx, *qy;
    const double *zb, *area, *phi, *beta, *cx, *cy;
    const double *fc1, *fc2, *fcos, *fsin;               // friction
    const double *hsing, *hk1, *hk2, *hcos, *hsin;       // singular head loss
    double *dzdx, *dzdy;                                 // MUSCL slopes
    // incidence lists in reference order
    const int *int_ptr, *int_inc;  // entries (edge << 1) | side
    const int *int_nbr;            // the other cell of each incidence
    const int *bnd_ptr, *bnd_inc;  // entries: renumbered boundary edge
    // internal edges (renumbered)
    const int *eL, *eR;
    const double *elen, *enx, *eny, *exm, *eym;
    double *flx;                   // [6*n_int]: side 0 {dh, dqx, dqy}, side 1 the same
    double *swp;                   // [2*n_int]: CFL swept term of side 0, side 1
    // boundary edges (renumbered)
    const double *blen, *bnx, *bny;
    const int *bser;
    const int *bcell;              // cell of each boundary edge
    const int *bnd_cells;          // cells with at least one boundary edge, ascending
    int n_bnd_cells;
    double *bF1;
    double *partial;               // [n_blocks of k_cfl]: per-block dt candidates
    // wind (model_wind smithbanke): 0-based station of each cell, record
    // times, Cartesian velocity per station ([station][record])
    const int *wcode;
    const double *wind_t, *wind_x, *wind_y;
    int wind_nt;
    // runoff / infiltration: 0-based rain station of each cell, record times,
    // rate per station ([station][record]), coefficient and Horton rate per cell
    const int *rcode;
    const double *rain_t, *rain_v, *rcoef, *infil;
    int rain_nt;
    // momentum diffusion: per internal edge length / |centre distance along
    // the normal|, diffused discharges, per-block parabolic dt candidates
    const double *pfac;
    double *pqx, *pqy, *ppart;
};

struct Opts {
    int scheme, topo, friction, head_loss, neg_det, wind;
    int runoff, rain_water_only, infil;  // SW2DG_RUNOFF_*, model_porosity_rainfall, SW2DG_INFIL_*
    int mdiff;           // momentum diffusion (velocity_independent)
    double parab_span;   // t_max - t_0: the first bound of the parabolic substep
    double h_min, phi_min, v_max, fr_max;
    double adj_cut;  // reverse sweep: adjoint zeroed at cells shallower than this (0: exact)
};

// Workgroup b is dispatched to XCD b % 8.  Give each XCD a contiguous range
// of logical blocks (bijection on [0, nb)).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
__device__ __forceinline__ int xcd_block() { return xcd_remap((int)blockIdx.x, (int)gridDim.x); }

inline int nblocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

// host-side device-order tables of one mesh
struct MeshTables {
    int64_t nc = 0, nb = 0, ni = 0;
    std::vector<int> perm, inv, eperm, bperm, binv;  // perm[new] = reference index
    std::vector<int> iptr, bptr, iinc, binc;
    std::vector<int> eL, eR;
    std::vector<double> elen, enx, eny, exm, eym;
    std::vector<double> blen, bnx, bny;
    std::vector<int> bser, bcell, bcells;
    int n_bcells = 0;
};
void build_tables(const sw2dg_mesh *mesh, MeshTables &T);

struct AssimState;  // 4D-Var records and buffers (sw2dg_adjoint.hip)
void assim_destroy(AssimState *a);

}  // namespace sw2dg_detail

struct sw2dg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    sw2dg_params prm{};
    int64_t nc = 0, nb = 0, ni = 0;
    std::vector<int> perm, inv;     // new cell -> old, old -> new
    std::vector<int> eperm, binv;   // new edge -> old edge ; old bnd -> new bnd
    std::vector<int> bperm;
    std::vector<int> bser_host;     // reference order
    sw2dg_detail::MeshDev m{};
    sw2dg_detail::Opts o{};
    sw2dg_detail::BcDev bc{};
    sw2dg_detail::Clock *ck = nullptr;
    sw2dg_detail::Clock hck{};
    std::vector<void *> allocs;
    int64_t bytes = 0;
    int n_series = 0, n_times = 0;
    size_t wind_nv = 0;  // station x record size of the device wind tables
    size_t rain_nv = 0;  // station x record size of the device rain table
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool state_set = false;
    std::vector<double> series_t_host, series_v_host;  // BC table (host copy)
    std::vector<int> series_type_host;
    std::vector<double> h0, qx0, qy0;                  // state given to set_state (new order)
    double t0 = 0.0;
    sw2dg_detail::AssimState *assim = nullptr;

    template <class T>
    T *dalloc(size_t n) {
        void *p = nullptr;
        const size_t b = std::max<size_t>(n, 1) * sizeof(T);
        if (hipMalloc(&p, b) != hipSuccess) return nullptr;
        if (hipMemset(p, 0, b) != hipSuccess) return nullptr;
        allocs.push_back(p);
        bytes += (int64_t)b;
        return (T *)p;
    }
    template <class T>
    T *upload(const std::vector<T> &v) {
        T *d = dalloc<T>(v.size());
        if (d && !v.empty() &&
            hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return nullptr;
        return d;
    }
    // reference order -> device order (cells)
    template <class T>
    std::vector<T> to_dev_cells(const T *src) const {
        std::vector<T> v(nc);
        for (int64_t k = 0; k < nc; ++k) v[k] = src[perm[k]];
        return v;
    }
};

namespace sw2dg_detail {
int fail(sw2dg_ctx *c, int code, const std::string &msg);
}

#define HIPCHK(c, x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return sw2dg_detail::fail((c), SW2DG_ERR_HIP,                                 \
                                      std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)
