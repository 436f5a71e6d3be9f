// emu_driver.cpp -- runs the search and strain kernels of dvc_kernels.hip on the CPU
// (see dvc_emu.h).  Test infrastructure: tests/test_emu_kernels.py compares these
// runs with the oracle, so kernel logic is checked without a GPU.
#define DVC_EMU 1
#include "dvc_kernels.hip"
#include "dvc_geometry.h"

using namespace dvc;

// Slice launches of a chain (SliceArgs) with their buffers (two of each, by launch parity).
struct EmuSlices {
    int nslice = 0, per = 0, evals = 0, cnt8 = 0;
    std::vector<char> sliced;
    std::vector<unsigned char> state;
    std::vector<double> part;
};

template <int BPV>
static void run_slice(const SearchArgs &a, const SliceArgs &sa, int threads) {
    const unsigned blocks = (unsigned)(sa.cnt8 * sa.nslice);
    const size_t slds = (size_t)(threads / 64) * 64 * 17 * 4;
    emu::grid_dim.x = blocks;
    const unsigned step = threads > 256 ? 2 : 4;
    for (unsigned b0 = 0; b0 < blocks; b0 += step) {
        const unsigned nb = blocks - b0 < step ? blocks - b0 : step;
        if (threads == 256) emu::run_grid(b0, nb, 256, slds, [&] { slice_kernel<BPV, 256, false>(a, sa); });
        else emu::run_grid(b0, nb, 512, slds, [&] { slice_kernel<BPV, 512, false>(a, sa); });
    }
}

// One launch per dependency level; the workgroups of a level run concurrently (eight at
// a time).  Levels marked in sl.sliced run as slice launches followed by the resuming
// one-workgroup search, as the library enqueues them.
template <int BPV, int B>
static void run_search(const SearchArgs &a, size_t lds, const std::vector<int> &lp, const std::vector<int> &lo,
                       EmuSlices *sl = nullptr) {
    for (size_t L = 0; L + 1 < lo.size(); L++) {
        int cnt = lo[L + 1] - lo[L];
        const int *pts = lp.data() + lo[L];
        SliceArgs rs{};
        if (sl && sl->sliced[L]) {
            SliceArgs sa{};
            sa.level_pts = pts;
            sa.cnt = cnt;
            sa.cnt8 = (cnt + 7) / 8 * 8;
            sa.nslice = sl->nslice;
            sa.per = sl->per;
            for (int q = 0; q < 2; q++) {
                sa.state[q] = sl->state.data() + (size_t)q * sl->cnt8 * kStateStride;
                sa.part[q] = sl->part.data() + (size_t)q * sl->cnt8 * sl->nslice * kPartStride;
            }
            const int threads = sl->per >= 512 ? 512 : 256;
            for (int t = 0; t < sl->evals; t++) {
                sa.t = t;
                run_slice<BPV>(a, sa, threads);
            }
            rs = sa;
            rs.t = sl->evals;
        }
        emu::grid_dim.x = cnt;
        for (int b0 = 0; b0 < cnt; b0 += 8) {
            int nb = cnt - b0 < 8 ? cnt - b0 : 8;
            if (rs.cnt > 0)
                emu::run_grid(b0, nb, 512, (size_t)8 * 64 * 17 * 4, [&] { resume_kernel<BPV, 512, false>(a, pts, rs); });
            else
                emu::run_grid(b0, nb, B, lds, [&] { search_kernel<BPV, B, false>(a, pts); });
        }
    }
}

// slice_wgs: -1 no slice launches, else the largest workgroup count of a sliced launch
// (0: every launch with at least two slices); slice_evals slice launches per chain
// launch; slice_per samples per slice.  Slice launches need precompute.
extern "C" int emu_search_sliced(const dvc_run_params *rc, const void *ref, const void *tar, int bpv, int W, int H,
                                 int T, const double *pts, int npts, const int *order, int nproc, const int *neigh,
                                 int k, int block_threads, int gather_mode, int precompute, int slice_wgs,
                                 int slice_evals, int slice_per, dvc_point_result *out, int *sliced_launches) {
    Geometry G = make_geometry(*rc);
    int N = npts;
    if (nproc > 1 && nproc < npts) N = nproc;
    std::vector<int> pos;
    if (!order_positions(order, npts, N, pos)) return -1;
    std::vector<double> sph;
    double omn[3] = {0, 0, 0}, omx[3] = {0, 0, 0};
    if (rc->geom == DVC_GEOM_SPHERE) {
        sphere_template(*rc, G, sph, omn, omx);
        locality_order(sph);
    }
    size_t vbytes = (size_t)W * H * T * bpv;
    std::vector<uint32_t> v0(vbytes / 4 + 32), v1(vbytes / 4 + 32);
    std::memcpy(v0.data(), ref, vbytes);
    std::memcpy(v1.data(), tar, vbytes);
    std::vector<dvc_point_result> res(npts);
    std::vector<int> lp, lo;
    dependency_levels(order, pos, neigh, npts, k, N, lp, lo);
    std::memset(res.data(), 0, sizeof(dvc_point_result) * npts);
    SearchArgs a;
    fill_search_args(a, *rc, G, bpv, W, H, T, npts, N, k, omn, omx);
    a.vol[0] = (const uint8_t *)v0.data();
    a.vol[1] = (const uint8_t *)v1.data();
    if (gather_mode == 0) a.gather32 = 0;          // force the 64-bit row addressing
    a.sph = sph.data(); a.pts = pts; a.order = order; a.pos = pos.data(); a.neigh = neigh;
    a.res = res.data();
    a.nbres = res.data();
    size_t lds = (size_t)a.nsamp_pad * 4 + (size_t)(block_threads / 64) * 64 * 17 * 4;   // samples + staging
    std::vector<float> rvals;
    std::vector<RefStats> rstat;
    if (precompute) {
        // reference sub-volumes of all searched points first (reference_kernel)
        rvals.assign((size_t)N * a.nsamp_pad, 0.0f);
        rstat.assign(N, RefStats());
        a.rvals = rvals.data();
        a.rstat = rstat.data();
        emu::grid_dim.x = N;
        for (int b0 = 0; b0 < N; b0 += 8) {
            int nb = N - b0 < 8 ? N - b0 : 8;
            if (bpv == 1 && block_threads == 64) emu::run_grid(b0, nb, 64, 0, [&] { reference_kernel<1, 64, false>(a); });
            else if (bpv == 1 && block_threads == 128) emu::run_grid(b0, nb, 128, 0, [&] { reference_kernel<1, 128, false>(a); });
            else if (bpv == 2 && block_threads == 64) emu::run_grid(b0, nb, 64, 0, [&] { reference_kernel<2, 64, false>(a); });
            else if (bpv == 2 && block_threads == 128) emu::run_grid(b0, nb, 128, 0, [&] { reference_kernel<2, 128, false>(a); });
            else return -1;
        }
    }
    EmuSlices sl;
    EmuSlices *slp = nullptr;
    if (sliced_launches) *sliced_launches = 0;
    if (slice_wgs >= 0) {
        if (!precompute || rc->basin_radius != 0.0 || (slice_per != 256 && slice_per % 512 != 0) || slice_evals < 1)
            return -1;
        sl.per = slice_per;
        sl.evals = slice_evals;
        sl.nslice = (a.nsamp + slice_per - 1) / slice_per;
        const std::vector<int> nsl(lo.size() > 0 ? lo.size() - 1 : 0, sl.nslice);
        const int cnt8 = slice_plan(lo, nsl, slice_wgs > 0 ? slice_wgs : 1 << 30, 8, sl.sliced);
        if (cnt8 > 0) {
            sl.cnt8 = cnt8;
            sl.state.assign((size_t)2 * cnt8 * kStateStride, 0xA5);
            sl.part.assign((size_t)2 * cnt8 * sl.nslice * kPartStride, -7.0);
            slp = &sl;
            int ns = 0;
            for (char c : sl.sliced) ns += c ? 1 : 0;
            if (sliced_launches) *sliced_launches = ns;
        }
    }
    if (bpv == 1 && block_threads == 64) run_search<1, 64>(a, lds, lp, lo, slp);
    else if (bpv == 1 && block_threads == 256) run_search<1, 256>(a, lds, lp, lo, slp);
    else if (bpv == 1 && block_threads == 128) run_search<1, 128>(a, lds, lp, lo, slp);
    else if (bpv == 2 && block_threads == 128) run_search<2, 128>(a, lds, lp, lo, slp);
    else if (bpv == 2 && block_threads == 64) run_search<2, 64>(a, lds, lp, lo, slp);
    else if (bpv == 2 && block_threads == 256) run_search<2, 256>(a, lds, lp, lo, slp);
    else return -1;
    for (int i = 0; i < npts; i++) {
        out[i] = res[i];
        if (pos[i] >= N) {
            out[i].status = DVC_NOT_SEARCH; out[i].iters = 0; out[i].obj = 0.0;
            for (int q = 0; q < 12; q++) out[i].par[q] = 0.0;
        }
    }
    return 0;
}

extern "C" int emu_search(const dvc_run_params *rc, const void *ref, const void *tar, int bpv, int W, int H, int T,
                          const double *pts, int npts, const int *order, int nproc, const int *neigh, int k,
                          int block_threads, int gather_mode, int precompute, dvc_point_result *out) {
    return emu_search_sliced(rc, ref, tar, bpv, W, H, T, pts, npts, order, nproc, neigh, k, block_threads, gather_mode,
                             precompute, -1, 0, 0, out, nullptr);
}

// gn_step<N> alone on one wave: with the SSD moments V[i][j] = A[i][j], V[i][12] = -b[i],
// V[i][13] = 0 it solves A x = b, by the column-pivoted QR of the search kernel (chol 0)
// or by its Cholesky solve with the QR fallback (chol 1), or by the lane-uniform LDL^T
// solve of the slice launches with the QR fallback (chol 2).
extern "C" int emu_gn_solve(int n, const double *A, const double *b, double *x, int chol) {
    if (n != 3 && n != 6 && n != 12) return -1;
    SearchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.obj = DVC_OBJ_SSD;
    a.nsamp = 1;
    std::vector<double> V(256, 0.0), ne(256, 0.0);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) V[i * 16 + j] = A[i * n + j];
        V[i * 16 + 12] = -b[i];
    }
    emu::grid_dim.x = 1;
    emu::run_grid(0, 1, 64, 0, [&] {
        DVC_SHARED(PtShared, S);
        if (threadIdx.x == 0)
            for (int q = 0; q < 12; q++) { S.par[q] = 0.0; S.upd[q] = 0.0; }
        emu::wave_sync();
        if (chol == 2) {
            if (n == 3) gn_step<3, true, true>(a, S, V.data(), ne.data());
            else if (n == 6) gn_step<6, true, true>(a, S, V.data(), ne.data());
            else gn_step<12, true, true>(a, S, V.data(), ne.data());
        } else if (chol) {
            if (n == 3) gn_step<3, true>(a, S, V.data(), ne.data());
            else if (n == 6) gn_step<6, true>(a, S, V.data(), ne.data());
            else gn_step<12, true>(a, S, V.data(), ne.data());
        } else {
            if (n == 3) gn_step<3, false>(a, S, V.data(), ne.data());
            else if (n == 6) gn_step<6, false>(a, S, V.data(), ne.data());
            else gn_step<12, false>(a, S, V.data(), ne.data());
        }
        emu::wave_sync();
        if (threadIdx.x == 0)
            for (int q = 0; q < n; q++) x[q] = S.par[q];
    });
    return 0;
}

extern "C" int emu_strain(const double *pts, const double *disp, const int *status, const double *objmin,
                          const int *neigh, int k, int n, int ndp, double objt, int refill, int yz, int zx, int xy,
                          double *out_disp, int *out_npts, double *out_swr, double *out_E, double *out_L,
                          double *out_DG) {
    StrainArgs a;
    a.pts = pts; a.disp = disp; a.objmin = objmin; a.status = status; a.neigh = neigh;
    a.k = k; a.n = n; a.ndp = ndp; a.refill = refill; a.yz = yz; a.zx = zx; a.xy = xy; a.objt = objt;
    a.out_disp = out_disp; a.out_swr = out_swr; a.out_E = out_E; a.out_L = out_L; a.out_DG = out_DG;
    a.out_npts = out_npts;
    int blocks = (n + kStrainWaves - 1) / kStrainWaves;
    emu::grid_dim.x = blocks;
    for (int b = 0; b < blocks; b++) emu::run_grid(b, 1, 64 * kStrainWaves, 0, [&] { strain_kernel(a); });
    return 0;
}

// The chain's launch plans (dvc_geometry.h): dependency levels (cap <= 0) or launches of
// at most cap points.  level_pts[N], level_off[N + 1]; returns the number of launches.
extern "C" int emu_schedule(const int *order, int N, const int *neigh, int k, int cap, int *level_pts,
                            int *level_off) {
    std::vector<int> pos(N);
    for (int i = 0; i < N; i++) pos[order[i]] = i;
    std::vector<int> lp, lo;
    const int d = cap > 0 ? capacity_schedule(order, pos, neigh, k, N, cap, lp, lo)
                          : dependency_levels(order, pos, neigh, N, k, N, lp, lo);
    std::copy(lp.begin(), lp.end(), level_pts);
    std::copy(lo.begin(), lo.end(), level_off);
    return d;
}

// The reference kernel's workgroup order (dvc_geometry.h): out[N].
extern "C" void emu_spatial_order(const double *points, const int *order, int N, int xcds, int *out) {
    std::vector<int> v;
    spatial_block_order(points, order, N, xcds, v);
    std::copy(v.begin(), v.end(), out);
}

