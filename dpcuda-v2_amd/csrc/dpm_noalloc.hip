// dpm_noalloc.hip -- disk forces, energies and the total potential energy of caller-owned
// device arrays, with no device allocation: the working memory (cell grid, particle lists,
// reduction partials) is the file-scope array `scratch`, resident on every device the
// library runs on until the process ends.
//
// What it computes is what a disk system with the Verlet method (cutoff rc) returns from
// dpm_neighbor_init + dpm_zero_forces + dpm_compute_forces + dpm_reduce(DPM_RED_PE):
// Disk::calculateForces (src/particles/disk/disk.cu:218-221) over the rows of
// kernelUpdateNeighborList (src/particles/base/kernels.cu:256-293, all pairs within rc in
// ascending j), then totalPotentialEnergy (src/particles/base/particle.cu:1078-1080) in the
// library's canonical summation order.  The pairs come from a cell grid (cells at least rc
// wide; each particle's nine cells merged in ascending index), so the pair order, and every
// result bit, equal the list-based path.
#include <mutex>

#include "dpm_kernels.hpp"

namespace dpm {

// the rebuild kernels of dpm_core.hip that the grid reuses (counting sort of the cells)
__global__ void k_scan_a(Dev d, long long nc, int gated);
__global__ void k_scan_b(Dev d, long long nb_blocks, int gated);
__global__ void k_scan_c(Dev d, long long nc, int gated);
__global__ void k_cell_scatter(Dev d, int gated);
__global__ void k_cell_sort(Dev d, long long nc, int gated);
__global__ void k_final_sum(const double* part, int64_t nc, double* out);

// ---- scratch layout -----------------------------------------------------------------------
constexpr int64_t NA_MAX_N = 1 << 20;      // particles per call
constexpr int64_t NA_MAX_CELLS = 1 << 20;  // grid cells
constexpr int64_t NA_MAX_ALL_PAIRS = 4096; // particles of a box too small for a grid
constexpr int64_t NA_SCAN_TILE = 1024;     // cells per scan block (k_scan_a/c)
constexpr int64_t NA_MAX_SB = NA_MAX_CELLS / NA_SCAN_TILE + 1;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_OFF_CNT = 0;
constexpr size_t NA_OFF_START = NA_OFF_CNT + na_align(sizeof(int) * NA_MAX_CELLS);
constexpr size_t NA_OFF_CUR = NA_OFF_START + na_align(sizeof(int) * (NA_MAX_CELLS + 1));
constexpr size_t NA_OFF_BSUM = NA_OFF_CUR + na_align(sizeof(int) * NA_MAX_CELLS);
constexpr size_t NA_OFF_CELL = NA_OFF_BSUM + na_align(sizeof(int) * (NA_MAX_SB + 1));
constexpr size_t NA_OFF_PERM = NA_OFF_CELL + na_align(sizeof(int) * NA_MAX_N);
constexpr size_t NA_OFF_PART = NA_OFF_PERM + na_align(sizeof(int) * NA_MAX_N);
constexpr size_t NA_SCRATCH_BYTES = NA_OFF_PART + na_align(sizeof(double) * (NA_MAX_N / 64 + 1));
// 20.1 MiB per device

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

// The device address of `scratch` on each device, looked up once per device (the name is a
// host-side stand-in: only this address may be handed to kernels or copies).
constexpr int NA_MAX_DEVICES = 64;
static void* g_scratch[NA_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    DPM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        set_error("dpm_disk_forces_noalloc: device index out of range");
        return -1;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        DPM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("dpm_disk_forces_noalloc: no device address for the scratch array");
            return -1;
        }
        g_scratch[dev] = p;
    }
    *out = (unsigned char*)g_scratch[dev];
    return 0;
}

// ---- kernels --------------------------------------------------------------------------------
struct NaGrid {
    int ncx, ncy;    // cells per side (1 x 1: every particle is a candidate of every other)
    double csx, csy; // cell sides
};

__global__ void k_na_fill(int* a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) a[i] = 0;
}

__device__ __forceinline__ int na_coord(double pos, double cs, int n) {
    // getPBCCellIndex, include/particles/base/kernels.cuh:285-301
    long long idx = (long long)floor(pos / cs);
    long long m = idx % n;
    return (int)(m < 0 ? m + n : m);
}

// cell of each particle + counting-sort histogram (non-finite positions go to cell 0)
__global__ void k_na_bin(const double* x, const double* y, int64_t n, NaGrid g, int* cell,
                         int* cnt) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    if (g.ncx > 1) {
        double xi = x[i], yi = y[i];
        if (isfinite(xi) && isfinite(yi) && fabs(xi) < 1e15 * g.csx && fabs(yi) < 1e15 * g.csy)
            c = na_coord(xi, g.csx, g.ncx) + g.ncx * na_coord(yi, g.csy, g.ncy);
    }
    cell[i] = c;
    atomicAdd(&cnt[c], 1);
}

// Forces, energies and the closing chunk partials of the canonical PE sum.  One thread per
// particle; the candidates of the nine cells around i's (every particle, for a 1 x 1 grid,
// which has no cell lists: perm == nullptr) are taken in ascending index, and a pair
// counts when its minimum-image separation is within the cutoff (isWithinCutoffSquared,
// kernels.cuh:249-260), exactly as the listed pairs of the Verlet rows.
__global__ __launch_bounds__(BLOCK) void k_na_forces(const double* x, const double* y,
                                                     const double* rad, int64_t n, NaGrid g,
                                                     PairParams p, double rc2, const int* cell,
                                                     const int* start, const int* perm,
                                                     double* fx, double* fy, double* pe,
                                                     double* part) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const double xi = x[i], yi = y[i], ri = rad[i];
        int s[9], en[9], h[9];
        const int nl = g.ncx > 1 ? 9 : 1;
        if (nl == 9) {
            const int cid = cell[i];
            const int cx = cid % g.ncx, cy = cid / g.ncx;
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy) {
                const int yy = cy + oy < 0 ? cy + oy + g.ncy : (cy + oy >= g.ncy ? cy + oy - g.ncy : cy + oy);
#pragma unroll
                for (int ox = -1; ox <= 1; ++ox) {
                    const int xx = cx + ox < 0 ? cx + ox + g.ncx : (cx + ox >= g.ncx ? cx + ox - g.ncx : cx + ox);
                    const int c = xx + yy * g.ncx;
                    const int t = (oy + 1) * 3 + (ox + 1);
                    s[t] = start[c];
                    en[t] = start[c + 1];
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) s[t] = en[t] = 0;
            en[0] = (int)n;
        }
        auto at = [perm](int k) { return perm ? perm[k] : k; };
#pragma unroll
        for (int t = 0; t < 9; ++t) h[t] = s[t] < en[t] ? at(s[t]) : INT_MAX;
        double ax = 0.0, ay = 0.0;
        for (;;) {
            int best = INT_MAX, bt = -1;
#pragma unroll
            for (int t = 0; t < 9; ++t)
                if (h[t] < best) {
                    best = h[t];
                    bt = t;
                }
            if (bt < 0) break;
#pragma unroll
            for (int t = 0; t < 9; ++t)
                if (t == bt) {
                    s[t]++;
                    h[t] = s[t] < en[t] ? at(s[t]) : INT_MAX;
                }
            if (best == i) continue;
            const double xj = x[best], yj = y[best];
            const double dx = pbc(xi, xj, p.Lx, p.invLx, p.halfLx, 0);
            const double dy = pbc(yi, yj, p.Ly, p.invLy, p.halfLy, 0);
            if (dx * dx + dy * dy < rc2) {
                double fpx, fpy;
                const double ie = pair_law(xi, yi, ri, xj, yj, rad[best], p, fpx, fpy);
                ax += fpx;
                ay += fpy;
                e += ie;
            }
        }
        fx[i] = 0.0 + ax;
        fy[i] = 0.0 + ay;
        e = 0.0 + e;
        pe[i] = e;
    }
    // chunk partial of the canonical sum (k_chunk_sum's tree over 64 consecutive particles)
    const double v = wave_tree_sum(i < n ? e : 0.0);
    const int64_t c = i / CHUNK;
    if ((threadIdx.x % WAVE) == 0 && c * CHUNK < n) part[c] = v;
}

}  // namespace dpm

using namespace dpm;

extern "C" {

int dpm_disk_forces_noalloc(const double* x, const double* y, const double* radius, int64_t n,
                            double lx, double ly, double e_c, double n_c, double cutoff,
                            double* fx, double* fy, double* pe, double* pe_total, void* stream) {
    if (n < 1 || n > NA_MAX_N) {
        set_error("dpm_disk_forces_noalloc: particle count must be in [1, 1048576] (the size "
                  "of the library's scratch array)");
        return -1;
    }
    if (!x || !y || !radius || !fx || !fy || !pe || !pe_total) {
        set_error("dpm_disk_forces_noalloc: null array");
        return -1;
    }
    if (!(lx > 0) || !(ly > 0) || !(cutoff > 0) || !(e_c > 0) || !(n_c > 0)) {
        set_error("dpm_disk_forces_noalloc: box sides, cutoff, e_c and n_c must be positive");
        return -1;
    }
    unsigned char* base = nullptr;
    if (scratch_of_current_device(&base)) return -1;
    hipStream_t st = (hipStream_t)stream;
    // the grid: cells at least cutoff (1 + 1e-9) wide, at most NA_MAX_CELLS and about two
    // per particle; under three a side, one cell holds every particle
    const double cw = cutoff * (1.0 + 1e-9);
    double fx_c = std::floor(lx / cw), fy_c = std::floor(ly / cw);
    const double cap = std::min<double>((double)NA_MAX_CELLS, std::max(9.0, 2.0 * (double)n));
    while (fx_c * fy_c > cap) {
        if (fx_c >= fy_c) fx_c -= 1.0;
        else fy_c -= 1.0;
    }
    NaGrid g{1, 1, lx, ly};
    if (fx_c >= 3 && fy_c >= 3) g = NaGrid{(int)fx_c, (int)fy_c, lx / fx_c, ly / fy_c};
    if (g.ncx == 1 && n > NA_MAX_ALL_PAIRS) {
        set_error("dpm_disk_forces_noalloc: a box under three cutoffs a side takes at most "
                  "4096 particles (every pair is tested)");
        return -1;
    }
    const int64_t nc = (int64_t)g.ncx * g.ncy;
    // carve the scratch array
    Dev d{};
    d.N = n;
    d.cell_cnt = (int*)(base + NA_OFF_CNT);
    d.cell_start = (int*)(base + NA_OFF_START);
    d.cell_cur = (int*)(base + NA_OFF_CUR);
    d.scan_bsum = (int*)(base + NA_OFF_BSUM);
    d.cell_tmp = (int*)(base + NA_OFF_CELL);
    d.perm = (int*)(base + NA_OFF_PERM);
    double* part = (double*)(base + NA_OFF_PART);
    // the pair law's parameters (pair_params without a system)
    PairParams p{};
    p.Lx = lx;
    p.Ly = ly;
    p.invLx = 1.0 / lx;
    p.invLy = 1.0 / ly;
    p.halfLx = 0.5 * lx;
    p.halfLy = 0.5 * ly;
    p.e_c = e_c;
    p.n_c = n_c;
    p.ncm1 = n_c - 1.0;
    p.two_nc = 2.0 * n_c;
    int ex;
    p.inv2nc = (std::frexp(p.two_nc, &ex) == 0.5) ? 1.0 / p.two_nc : 0.0;
    p.fast = 0;
    p.fastdiv = 0;
    const int64_t nsb = (nc + NA_SCAN_TILE - 1) / NA_SCAN_TILE;
    const bool grid = g.ncx > 1;
    if (grid) {  // counting sort into the cells (cells cleared by a kernel: no memset node,
                 // so a captured graph replays correctly)
        hipLaunchKernelGGL(k_na_fill, pgrid(nc), dim3(BLOCK), 0, st, d.cell_cnt, nc);
        hipLaunchKernelGGL(k_na_bin, pgrid(n), dim3(BLOCK), 0, st, x, y, n, g, d.cell_tmp,
                           d.cell_cnt);
        hipLaunchKernelGGL(k_scan_a, dim3((unsigned)nsb), dim3(BLOCK), 0, st, d, (long long)nc, 0);
        hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, st, d, (long long)nsb, 0);
        hipLaunchKernelGGL(k_scan_c, dim3((unsigned)nsb), dim3(BLOCK), 0, st, d, (long long)nc, 0);
        hipLaunchKernelGGL(k_cell_scatter, pgrid(n), dim3(BLOCK), 0, st, d, 0);
        hipLaunchKernelGGL(k_cell_sort, pgrid(nc), dim3(BLOCK), 0, st, d, (long long)nc, 0);
    }
    hipLaunchKernelGGL(k_na_forces, pgrid(n), dim3(BLOCK), 0, st, x, y, radius, n, g, p,
                       cutoff * cutoff, d.cell_tmp, d.cell_start, grid ? d.perm : nullptr, fx,
                       fy, pe, part);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(WAVE), 0, st, part, (n + CHUNK - 1) / CHUNK,
                       pe_total);
    return check_launch();
}

}  // extern "C"
