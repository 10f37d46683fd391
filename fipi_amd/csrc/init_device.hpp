// Seeded initial state on the device: the same values as the host
// generators of init.cpp (fipi_init_phi / fipi_init_uniform /
// fipi_init_rsa), computed on the GPU.  Included at the end of fipi.hip.
//
// Random sequential addition as a stream of candidates.  The host RSA draws
// candidate k at counters 3k, 3k+1, 3k+2 and accepts it when it keeps the
// minimum distance to every sphere placed before it; sphere i fails when
// max_attempts candidates in a row after the previous acceptance are
// rejected.  The device does the same over chunks of candidates:
//   1. every candidate of the chunk is tested against the spheres already
//      placed (a cell grid of cell >= the minimum distance); a conflict
//      there is final, since the placed set only grows;
//   2. the survivors are resolved among themselves in candidate order --
//      candidate q is accepted when every earlier survivor within the
//      minimum distance was rejected, and rejected when one of them was
//      accepted -- by rounds of one kernel until nothing is undecided (the
//      greedy order is the host's, so the decisions are the host's);
//   3. the host walks the accepted candidates in order, applies n and the
//      max_attempts rule, and the kept ones are placed.
// The result does not depend on the chunk size, which adapts so that about
// one survivor in four cells is expected (a cell overflowing its survivor
// capacity halves the chunk and redoes it).
#pragma once

namespace {

constexpr int RSA_CAP = 32;        // placed spheres per grid cell (spaced by >= the minimum distance)
constexpr int RSA_CHUNK_CAP = 64;  // survivors per grid cell within one chunk

// uniform double in [0,1) for (seed, counter), as init.cpp's u01
__device__ __forceinline__ double init_u01(uint64_t seed, uint64_t ctr)
{
    const uint64_t h = splitmix64(splitmix64(seed) ^ (ctr * 0xD1B54A32D192ED03ull));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// out[i] = a + b * (2u - 1) (mode 0, phi) or a + (b - a) * u (mode 1)
__global__ void k_init_values(double *out, int64_t n, double a, double b, uint64_t seed, int mode)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = init_u01(seed, (uint64_t)i);
    out[i] = mode == 0 ? a + b * (2.0 * u - 1.0) : a + (b - a) * u;
}

struct RsaGeom {
    double L[3], cs[3], dmin2;
    int nc[3], r[3];  // cells per axis and the neighbour reach (1, or nc-1 below 3 cells)
};

__device__ __forceinline__ double rsa_dist2(const RsaGeom &g, const double *p, const double *q)
{
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double d = p[a] - q[a];
        d -= g.L[a] * rint(d / g.L[a]);
        s += d * d;
    }
    return s;
}

__device__ __forceinline__ int rsa_cell(const RsaGeom &g, const double *p, int *ci)
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        ci[a] = (int)floor(p[a] / g.cs[a]);
        if (ci[a] >= g.nc[a]) ci[a] = g.nc[a] - 1;
    }
    return ci[0] + g.nc[0] * (ci[1] + g.nc[1] * ci[2]);
}

// visit the cells around ci (the host's neighbour loops, duplicates included
// when an axis has fewer than 3 cells)
template <class F>
__device__ __forceinline__ bool rsa_any_neighbour(const RsaGeom &g, const int *ci, F &&f)
{
    for (int dz = -g.r[2]; dz <= g.r[2]; dz++)
        for (int dy = -g.r[1]; dy <= g.r[1]; dy++)
            for (int dx = -g.r[0]; dx <= g.r[0]; dx++) {
                const int cx = (ci[0] + dx + g.nc[0]) % g.nc[0], cy = (ci[1] + dy + g.nc[1]) % g.nc[1],
                          cz = (ci[2] + dz + g.nc[2]) % g.nc[2];
                if (f(cx + g.nc[0] * (cy + g.nc[1] * cz))) return true;
            }
    return false;
}

// step 1: candidates k0 .. k0+C-1 against the placed spheres; survivors go
// into the chunk grid (status 0 = undecided, 2 = rejected)
__global__ void k_rsa_candidates(RsaGeom g, uint64_t seed, int64_t k0, int C, const double *__restrict__ pos,
                                 const int *__restrict__ pcnt, const int *__restrict__ pids, double *__restrict__ cpos,
                                 int *__restrict__ status, int *__restrict__ ccnt, int *__restrict__ cids,
                                 int *__restrict__ overflow)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const uint64_t ctr = 3 * (uint64_t)(k0 + q);
    double p[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        p[a] = g.L[a] * init_u01(seed, ctr + a);
        if (p[a] >= g.L[a]) p[a] = 0.0;
        cpos[3 * (size_t)q + a] = p[a];
    }
    int ci[3];
    const int c = rsa_cell(g, p, ci);
    const bool hit = rsa_any_neighbour(g, ci, [&](int cell) {
        const int m = min(pcnt[cell], RSA_CAP);
        for (int s = 0; s < m; s++)
            if (rsa_dist2(g, p, pos + 3 * (size_t)pids[(size_t)cell * RSA_CAP + s]) < g.dmin2) return true;
        return false;
    });
    status[q] = hit ? 2 : 0;
    if (!hit) {
        const int slot = atomicAdd(&ccnt[c], 1);
        if (slot < RSA_CHUNK_CAP) cids[(size_t)c * RSA_CHUNK_CAP + slot] = q;
        else atomicOr(overflow, 1);
    }
}

// step 2, one round: undecided q becomes 1 (accepted) when every earlier
// survivor within the minimum distance is rejected, 2 when one is accepted
__global__ void k_rsa_resolve(RsaGeom g, int C, const double *__restrict__ cpos, int *status,
                              const int *__restrict__ ccnt, const int *__restrict__ cids, int *__restrict__ undecided)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C || status[q] != 0) return;
    const double *p = cpos + 3 * (size_t)q;
    int ci[3];
    rsa_cell(g, p, ci);
    bool pending = false;
    const bool rejected = rsa_any_neighbour(g, ci, [&](int cell) {
        const int m = min(ccnt[cell], RSA_CHUNK_CAP);
        for (int s = 0; s < m; s++) {
            const int j = cids[(size_t)cell * RSA_CHUNK_CAP + s];
            if (j >= q || rsa_dist2(g, p, cpos + 3 * (size_t)j) >= g.dmin2) continue;
            const int st = __atomic_load_n(&status[j], __ATOMIC_RELAXED);
            if (st == 1) return true;
            if (st == 0) pending = true;
        }
        return false;
    });
    if (rejected) status[q] = 2;
    else if (!pending) status[q] = 1;
    else atomicAdd(undecided, 1);
}

// step 3: place the kept candidates keep[0 .. m) as spheres base .. base+m
__global__ void k_rsa_place(RsaGeom g, const int *__restrict__ keep, int m, int64_t base,
                            const double *__restrict__ cpos, double *__restrict__ pos, int *__restrict__ pcnt,
                            int *__restrict__ pids, int *__restrict__ overflow)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double *p = cpos + 3 * (size_t)keep[i];
    const int64_t s = base + i;
#pragma unroll
    for (int a = 0; a < 3; a++) pos[3 * s + a] = p[a];
    int ci[3];
    const int c = rsa_cell(g, p, ci);
    const int slot = atomicAdd(&pcnt[c], 1);
    if (slot < RSA_CAP) pids[(size_t)c * RSA_CAP + slot] = (int)s;
    else atomicOr(overflow, 1);
}

// fill an int array
__global__ void k_init_fill_int(int *a, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// device allocations of one RSA run, freed on every return path
struct InitMem {
    std::vector<void *> mem;
    hipError_t alloc(void **p, size_t b)
    {
        const hipError_t e = hipMalloc(p, b ? b : 16);
        if (e == hipSuccess) mem.push_back(*p);
        return e;
    }
    ~InitMem()
    {
        for (void *p : mem) (void)hipFree(p);
    }
};

thread_local std::string g_init_error;

int64_t init_fail(int code, const std::string &msg)
{
    g_init_error = msg;
    return code;
}

}  // namespace

extern "C" {

int fipi_init_phi_device(int64_t g, double mean, double amp, uint64_t seed, double *d_phi, void *stream)
{
    if (g < 0 || (g > 0 && !d_phi)) return FIPI_ERR_ARG;
    if (g == 0) return FIPI_OK;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_init_values, dim3((unsigned)((g + 255) / 256)), dim3(256), 0, st, d_phi, g, mean, amp, seed, 0);
    return hipStreamSynchronize(st) == hipSuccess && hipGetLastError() == hipSuccess ? FIPI_OK : FIPI_ERR_HIP;
}

int fipi_init_uniform_device(int64_t n, double lo, double hi, uint64_t seed, double *d_out, void *stream)
{
    if (n < 0 || (n > 0 && !d_out)) return FIPI_ERR_ARG;
    if (n == 0) return FIPI_OK;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_init_values, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_out, n, lo, hi, seed, 1);
    return hipStreamSynchronize(st) == hipSuccess && hipGetLastError() == hipSuccess ? FIPI_OK : FIPI_ERR_HIP;
}

int64_t fipi_init_rsa_device(int64_t n, double r, double gap, double lx, double ly, double lz, uint64_t seed,
                             int64_t max_attempts, double *d_pos, void *stream)
{
    if (n < 0 || (n > 0 && !d_pos) || !(r > 0) || !(2.0 * r + gap > 0) || !(lx > 0 && ly > 0 && lz > 0) ||
        max_attempts < 1)
        return init_fail(FIPI_ERR_ARG, "invalid argument");
    if (n > INT32_MAX / 2) return init_fail(FIPI_ERR_ARG, "n too large");
    if (n == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const double dmin = 2.0 * r + gap;
    RsaGeom g{};
    g.L[0] = lx;
    g.L[1] = ly;
    g.L[2] = lz;
    g.dmin2 = dmin * dmin;
    int64_t ncell = 1;
    // cells of at least the minimum distance, as the host's (any such grid
    // finds the same conflicts), at most 128 a side
    for (int a = 0; a < 3; a++) {
        g.nc[a] = (int)std::min(std::floor(g.L[a] / dmin), 128.0);
        if (g.nc[a] < 1) g.nc[a] = 1;
        g.cs[a] = g.L[a] / g.nc[a];
        g.r[a] = g.nc[a] >= 3 ? 1 : g.nc[a] - 1;
        ncell *= g.nc[a];
    }
    const int64_t cmax = std::max<int64_t>(ncell / 4, 1024) * 16;  // largest chunk
    InitMem mm;
    int *pcnt, *pids, *ccnt, *cids, *status, *keep, *flags;
    double *cpos;
    if (mm.alloc((void **)&pcnt, sizeof(int) * ncell) || mm.alloc((void **)&pids, sizeof(int) * ncell * RSA_CAP) ||
        mm.alloc((void **)&ccnt, sizeof(int) * ncell) ||
        mm.alloc((void **)&cids, sizeof(int) * ncell * RSA_CHUNK_CAP) ||
        mm.alloc((void **)&status, sizeof(int) * cmax) || mm.alloc((void **)&keep, sizeof(int) * cmax) ||
        mm.alloc((void **)&cpos, sizeof(double) * 3 * cmax) || mm.alloc((void **)&flags, sizeof(int) * 2))
        return init_fail(FIPI_ERR_HIP, "hipMalloc failed");
    auto ok = [&](hipError_t e) { return e == hipSuccess; };
    auto launch_ok = [&]() { return ok(hipGetLastError()); };
    hipLaunchKernelGGL(k_init_fill_int, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, pcnt, ncell, 0);
    if (!launch_ok()) return init_fail(FIPI_ERR_HIP, "launch failed");
    std::vector<int> hstat;
    std::vector<int> hkeep;
    int64_t placed = 0, k_last = -1, k0 = 0;
    int64_t C = std::min<int64_t>(std::max<int64_t>(ncell / 4, 64), cmax);
    const int64_t target = std::max<int64_t>(ncell / 4, 64);  // survivors per chunk aimed at
    while (placed < n) {
        // candidates k0 .. k0+C-1 and their survivors
        hipLaunchKernelGGL(k_init_fill_int, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, ccnt, ncell, 0);
        hipLaunchKernelGGL(k_init_fill_int, dim3(1), dim3(64), 0, st, flags, 2, 0);
        hipLaunchKernelGGL(k_rsa_candidates, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, g, seed, k0, (int)C,
                           (const double *)d_pos, (const int *)pcnt, (const int *)pids, cpos, status, ccnt, cids,
                           flags);
        int hf[2];
        if (!launch_ok() || !ok(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, st)) ||
            !ok(hipStreamSynchronize(st)))
            return init_fail(FIPI_ERR_HIP, "RSA candidates failed");
        if (hf[0]) {
            // a cell got more survivors than it holds: the same candidates in
            // smaller chunks (the decisions do not depend on the chunk)
            if (C <= 64) return init_fail(FIPI_ERR_OVERFLOW, "RSA chunk grid overflow");
            C /= 2;
            continue;
        }
        // resolve the survivors among themselves
        for (int round = 0;; round++) {
            hipLaunchKernelGGL(k_init_fill_int, dim3(1), dim3(64), 0, st, flags + 1, 1, 0);
            hipLaunchKernelGGL(k_rsa_resolve, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, g, (int)C,
                               (const double *)cpos, status, (const int *)ccnt, (const int *)cids, flags + 1);
            int und = 0;
            if (!launch_ok() || !ok(hipMemcpyAsync(&und, flags + 1, sizeof(int), hipMemcpyDeviceToHost, st)) ||
                !ok(hipStreamSynchronize(st)))
                return init_fail(FIPI_ERR_HIP, "RSA resolve failed");
            if (und == 0) break;
            // each round decides at least the earliest undecided survivor
            if (round > C) return init_fail(FIPI_ERR_HIP, "RSA resolve did not converge");
        }
        hstat.resize(C);
        if (!ok(hipMemcpyAsync(hstat.data(), status, sizeof(int) * C, hipMemcpyDeviceToHost, st)) ||
            !ok(hipStreamSynchronize(st)))
            return init_fail(FIPI_ERR_HIP, "RSA status copy failed");
        // the host's order: n spheres, max_attempts misses in a row end it
        hkeep.clear();
        int64_t survivors = 0;
        bool jammed = false;
        for (int64_t q = 0; q < C && placed + (int64_t)hkeep.size() < n; q++) {
            if (hstat[q] != 2) survivors++;
            if (hstat[q] != 1) continue;
            const int64_t k = k0 + q;
            if (k - k_last - 1 >= max_attempts) {
                jammed = true;
                break;
            }
            hkeep.push_back((int)q);
            k_last = k;
        }
        if (!jammed && placed + (int64_t)hkeep.size() < n && (k0 + C - 1) - k_last >= max_attempts) jammed = true;
        const int m = (int)hkeep.size();
        if (m > 0) {
            if (!ok(hipMemcpyAsync(keep, hkeep.data(), sizeof(int) * m, hipMemcpyHostToDevice, st)))
                return init_fail(FIPI_ERR_HIP, "RSA keep copy failed");
            hipLaunchKernelGGL(k_init_fill_int, dim3(1), dim3(64), 0, st, flags, 1, 0);
            hipLaunchKernelGGL(k_rsa_place, dim3((m + 255) / 256), dim3(256), 0, st, g, (const int *)keep, m, placed,
                               (const double *)cpos, d_pos, pcnt, pids, flags);
            int of = 0;
            if (!launch_ok() || !ok(hipMemcpyAsync(&of, flags, sizeof(int), hipMemcpyDeviceToHost, st)) ||
                !ok(hipStreamSynchronize(st)))
                return init_fail(FIPI_ERR_HIP, "RSA place failed");
            if (of) return init_fail(FIPI_ERR_OVERFLOW, "RSA grid cell overflow");
            placed += m;
        }
        if (jammed) break;
        k0 += C;
        // next chunk: about `target` survivors
        const double rate = (double)std::max<int64_t>(survivors, 1) / (double)C;
        C = std::min<int64_t>(std::max<int64_t>((int64_t)(target / rate), 64), cmax);
    }
    return placed;
}

const char *fipi_init_last_error(void) { return g_init_error.c_str(); }

}  // extern "C"
