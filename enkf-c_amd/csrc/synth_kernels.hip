// synth_kernels.hip -- synthetic inputs for the benchmark and the tests
// (libenkfsynth.so; not part of the engine library libenkfg.so).
//
// Generates, on the device and without any other GPU library, the pieces of
// a synthetic analysis problem: Gaussian random fields, the ensemble of a
// tile in the engine's device layout [field][cell][member], and the
// ensemble observation anomalies HA [member][obs] with innovations.  The
// random numbers are Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011), keyed by a seed and counted
// by (stream, index), so every value is reproducible from its coordinates.
//
// Random fields use random Fourier features (Rahimi and Recht, NIPS 2007):
// f(x) = sqrt(2/K) sum_k cos(w_k . x + phi_k) with w_k ~ N(0, I / L^2)
// (rad/km) and phi_k ~ U(0, 2 pi), whose covariance is exp(-r^2 / (2 L^2));
// each field is then shifted and scaled to zero mean and unit variance over
// the grid.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds: counter c, key k
__host__ __device__ inline U4 philox(U4 c, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c.x, hi0, lo0);
        mulhilo(0xCD9E8D57u, c.z, hi1, lo1);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// uniform in (0, 1)
__device__ inline float u01(uint32_t u)
{
    return ((float)(u >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

__device__ inline double u01d(uint32_t a, uint32_t b)
{
    const uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    return ((double)(v & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
}

constexpr int kFeat = 256;   // Fourier features per field
constexpr double kTwoPi = 6.283185307179586;

// features of field f: [f][k] = (wx, wy, phi, 0)
__global__ void grf_features_kernel(float4* feat, int n, double corr_km, uint32_t seed)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * kFeat)
        return;
    const U4 r = philox(U4{(uint32_t)idx, 0x5EEDu, 0u, 0u}, seed, 0x9F1Du);
    // Box-Muller for the two frequency components
    const float a = u01(r.x), b = u01(r.y);
    const float rad = sqrtf(-2.0f * logf(a)) / (float)corr_km;
    float s, c;
    sincosf((float)kTwoPi * b, &s, &c);
    feat[idx] = make_float4(rad * c, rad * s, (float)kTwoPi * u01(r.z), 0.0f);
}

// out[f][j][i] = sqrt(2/K) sum_k cos(wx x + wy y + phi), x = i dx, y = j dx
__global__ void __launch_bounds__(256) grf_eval_kernel(const float4* __restrict__ feat, float* __restrict__ out,
                                                       int nj, int ni, float dx_km)
{
    __shared__ float4 F[kFeat];
    const int f = blockIdx.y;
    for (int k = threadIdx.x; k < kFeat; k += blockDim.x)
        F[k] = feat[(size_t)f * kFeat + k];
    __syncthreads();
    const long long ncell = (long long)nj * ni;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell)
        return;
    const float x = (float)(cell % ni) * dx_km;
    const float y = (float)(cell / ni) * dx_km;
    float acc = 0.0f;
#pragma unroll 8
    for (int k = 0; k < kFeat; ++k)
        acc += __cosf(F[k].x * x + F[k].y * y + F[k].z);
    out[(size_t)f * ncell + cell] = acc * sqrtf(2.0f / (float)kFeat);
}

// per-field sum and sum of squares (double atomics, one pair per block)
__global__ void __launch_bounds__(256) moments_kernel(const float* __restrict__ v, long long ncell,
                                                      double* __restrict__ mom)
{
    const int f = blockIdx.y;
    double s = 0.0, q = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        const double x = v[(size_t)f * ncell + c];
        s += x;
        q += x * x;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_down(s, o);
        q += __shfl_down(q, o);
    }
    __shared__ double ps[4], pq[4];
    const int wv = threadIdx.x / 64, ln = threadIdx.x % 64;
    if (ln == 0) {
        ps[wv] = s;
        pq[wv] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) {
            s += ps[k];
            q += pq[k];
        }
        atomicAdd(&mom[2 * f], s);
        atomicAdd(&mom[2 * f + 1], q);
    }
}

__global__ void normalise_kernel(float* v, long long ncell, const double* __restrict__ mom)
{
    const int f = blockIdx.y;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell)
        return;
    const double mean = mom[2 * f] / (double)ncell;
    const double var = mom[2 * f + 1] / (double)ncell - mean * mean;
    const double sd = var > 0.0 ? sqrt(var) : 1.0;
    float* p = v + (size_t)f * ncell + c;
    *p = (float)(((double)*p - mean) / sd);
}

constexpr int kCells = 32;   // cells per block of tile_ens_kernel (LDS: 3 m kCells floats)

struct TileArgs {
    int nmem, nk;
    long long ncell;
    const float* Gt;     // [m][ncell] temp base fields
    const float* Gs;     // salt
    const float* Ge;     // eta
    const float* G2;     // [ncell] mean-state field
    const float* prof;   // [nk] anomaly amplitude by level
    float sd_t, sd_s, sd_e;
    float mean_t, mean_s, mean_e;
    float* E;            // [2 nk + 1][ncell][m]
};

// One block per kCells cells: the members of the base fields are staged in LDS
// (coalesced over cells), then every field of those cells is written with
// members contiguous.  Field f < nk: temp level f; nk <= f < 2 nk: salt level
// f - nk; f = 2 nk: eta.
//   E = mean - 0.2 k + 0.3 G2 + sd prof[k] (cos(0.05 k) G[e] + sin(0.05 k) G[e-1])
//   E_eta = mean_e + 0.2 G2 + sd_e G_e[e]
__global__ void __launch_bounds__(256) tile_ens_kernel(TileArgs a)
{
    extern __shared__ float lds[];   // [3][m][kCells]
    const int m = a.nmem;
    const long long c0 = (long long)blockIdx.x * kCells;
    const int nc = (int)min((long long)kCells, a.ncell - c0);
    for (int idx = threadIdx.x; idx < 3 * m * kCells; idx += blockDim.x) {
        const int v = idx / (m * kCells), e = (idx / kCells) % m, cc = idx % kCells;
        const float* G = v == 0 ? a.Gt : (v == 1 ? a.Gs : a.Ge);
        lds[idx] = cc < nc ? G[(size_t)e * a.ncell + c0 + cc] : 0.0f;
    }
    __syncthreads();
    const int nfield = 2 * a.nk + 1;
    for (int f = 0; f < nfield; ++f) {
        const int v = f < a.nk ? 0 : (f < 2 * a.nk ? 1 : 2);
        const int k = v == 2 ? 0 : f - v * a.nk;
        const float mean = v == 0 ? a.mean_t : (v == 1 ? a.mean_s : a.mean_e);
        const float sd = v == 0 ? a.sd_t : (v == 1 ? a.sd_s : a.sd_e);
        const float amp = v == 2 ? sd : sd * a.prof[k];
        float sk, ck;
        sincosf(0.05f * (float)k, &sk, &ck);
        const float* L = lds + (size_t)v * m * kCells;
        for (int idx = threadIdx.x; idx < nc * m; idx += blockDim.x) {
            const int cc = idx / m, e = idx % m;
            const float g2 = a.G2[c0 + cc];
            float val;
            if (v == 2) {
                val = mean + 0.2f * g2 + amp * L[e * kCells + cc];
            } else {
                const int ep = e == 0 ? m - 1 : e - 1;
                val = mean - 0.2f * (float)k + 0.3f * g2 + amp * (ck * L[e * kCells + cc] + sk * L[ep * kCells + cc]);
            }
            a.E[((size_t)f * a.ncell + c0 + cc) * m + e] = val;
        }
    }
}

struct HaArgs {
    int nmem, nobs;
    long long ncell;
    const float* G[3];     // base fields per variable [m][ncell]: temp, salt, eta
    const int* cell;       // [nobs]
    const int* var;        // [nobs] 0 temp, 1 salt, 2 eta
    const int* lev;        // [nobs]
    const float* prof;     // [nk]
    const float* estd;     // [nobs]
    float sd[3];
    uint32_t seed;
    float* HA;             // [m][nobs]
    double* innov;         // [nobs]
};

// HA[e][o] = sd G_var[e][cell_o] (prof[k_o] for 3D variables), centred over
// the members in FP32; innov_o = N(0, 1) sqrt(spread_o^2 + estd_o^2) with the
// members' sample spread (unbiased) in FP64.
__global__ void __launch_bounds__(256) tile_ha_kernel(HaArgs a)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= a.nobs)
        return;
    const int v = a.var[o];
    const float* G = a.G[v];
    const long long c = a.cell[o];
    const float amp = v == 2 ? a.sd[2] : a.sd[v] * a.prof[a.lev[o]];
    float sum = 0.0f;
    for (int e = 0; e < a.nmem; ++e)
        sum += amp * G[(size_t)e * a.ncell + c];
    const float mean = sum / (float)a.nmem;
    double q = 0.0;
    for (int e = 0; e < a.nmem; ++e) {
        const float x = amp * G[(size_t)e * a.ncell + c] - mean;
        a.HA[(size_t)e * a.nobs + o] = x;
        q += (double)x * (double)x;
    }
    const double spread2 = a.nmem > 1 ? q / (double)(a.nmem - 1) : 0.0;
    const U4 r = philox(U4{(uint32_t)o, 0x1A0Fu, 0u, 0u}, a.seed, 0x77u);
    const double u1 = u01d(r.x, r.y), u2 = u01d(r.z, r.w);
    const double z = sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2);
    const double es = a.estd[o];
    a.innov[o] = z * sqrt(spread2 + es * es);
}

// uniform and normal samples on the host side of the generator (the same
// Philox stream the kernels use), for host-side positions
__global__ void uniform_kernel(double* out, long long n, uint32_t seed, uint32_t stream)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const U4 r = philox(U4{(uint32_t)i, (uint32_t)(i >> 32), stream, 0x0u}, seed, 0x3C6Eu);
    out[i] = u01d(r.x, r.y);
}

// a plain copy, 16 bytes per thread and step: the reference reading of the
// hardware byte counters (n16 = bytes / 16)
__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, long long n16)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int ret(hipError_t e)
{
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace

extern "C" {

int es_malloc(void** p, size_t bytes)
{
    return ret(hipMalloc(p, bytes > 0 ? bytes : 1));
}

int es_free(void* p)
{
    return ret(hipFree(p));
}

int es_h2d(void* dst, const void* src, size_t bytes)
{
    return ret(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int es_d2h(void* dst, const void* src, size_t bytes)
{
    return ret(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int es_zero(void* dst, size_t bytes)
{
    return ret(hipMemset(dst, 0, bytes));
}

int es_sync(void)
{
    return ret(hipDeviceSynchronize());
}

const char* es_error_string(int code)
{
    return hipGetErrorString((hipError_t)code);
}

// n unit-variance random fields [n][nj][ni] (device out), spacing dx_km,
// correlation length corr_km, from seed
int es_grf(float* out, int n, int nj, int ni, double dx_km, double corr_km, unsigned int seed)
{
    if (n <= 0 || nj <= 0 || ni <= 0)
        return (int)hipErrorInvalidValue;
    const long long ncell = (long long)nj * ni;
    float4* feat = nullptr;
    double* mom = nullptr;
    hipError_t e = hipMalloc(&feat, (size_t)n * kFeat * sizeof(float4));
    if (e == hipSuccess)
        e = hipMalloc(&mom, (size_t)2 * n * sizeof(double));
    if (e == hipSuccess)
        e = hipMemset(mom, 0, (size_t)2 * n * sizeof(double));
    if (e == hipSuccess) {
        grf_features_kernel<<<(n * kFeat + 255) / 256, 256>>>(feat, n, corr_km, seed);
        e = hipGetLastError();
    }
    const dim3 grid((unsigned)((ncell + 255) / 256), (unsigned)n);
    if (e == hipSuccess) {
        grf_eval_kernel<<<grid, 256>>>(feat, out, nj, ni, (float)dx_km);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        moments_kernel<<<dim3(256, (unsigned)n), 256>>>(out, ncell, mom);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        normalise_kernel<<<grid, 256>>>(out, ncell, mom);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (feat != nullptr)
        (void)hipFree(feat);
    if (mom != nullptr)
        (void)hipFree(mom);
    return ret(e);
}

// the tile ensemble E [2 nk + 1][ncell][nmem] (device) from the base fields
int es_tile_ensemble(float* E, int nmem, int nk, long long ncell, const float* Gt, const float* Gs, const float* Ge,
                     const float* G2, const float* prof, const float* sd, const float* mean)
{
    if (nmem <= 0 || nmem > 192 || nk <= 0 || ncell <= 0)
        return (int)hipErrorInvalidValue;
    TileArgs a{};
    a.nmem = nmem;
    a.nk = nk;
    a.ncell = ncell;
    a.Gt = Gt;
    a.Gs = Gs;
    a.Ge = Ge;
    a.G2 = G2;
    a.prof = prof;
    a.sd_t = sd[0];
    a.sd_s = sd[1];
    a.sd_e = sd[2];
    a.mean_t = mean[0];
    a.mean_s = mean[1];
    a.mean_e = mean[2];
    a.E = E;
    const size_t lds = (size_t)3 * nmem * kCells * sizeof(float);
    tile_ens_kernel<<<(unsigned)((ncell + kCells - 1) / kCells), 256, lds>>>(a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    return ret(e);
}

// HA [nmem][nobs] and innovations [nobs] (device) at the obs cells
int es_tile_ha(float* HA, double* innov, int nmem, int nobs, long long ncell, const float* Gt, const float* Gs,
               const float* Ge, const int* cell, const int* var, const int* lev, const float* prof, const float* estd,
               const float* sd, unsigned int seed)
{
    if (nmem <= 0 || nobs < 0 || ncell <= 0)
        return (int)hipErrorInvalidValue;
    if (nobs == 0)
        return 0;
    HaArgs a{};
    a.nmem = nmem;
    a.nobs = nobs;
    a.ncell = ncell;
    a.G[0] = Gt;
    a.G[1] = Gs;
    a.G[2] = Ge;
    a.cell = cell;
    a.var = var;
    a.lev = lev;
    a.prof = prof;
    a.estd = estd;
    a.sd[0] = sd[0];
    a.sd[1] = sd[1];
    a.sd[2] = sd[2];
    a.seed = seed;
    a.HA = HA;
    a.innov = innov;
    tile_ha_kernel<<<(nobs + 255) / 256, 256>>>(a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    return ret(e);
}

// copies `bytes` (a multiple of 16) from src to dst with copy_kernel
int es_copy(void* dst, const void* src, size_t bytes)
{
    if (bytes % 16 != 0)
        return (int)hipErrorInvalidValue;
    const long long n16 = (long long)(bytes / 16);
    if (n16 == 0)
        return 0;
    copy_kernel<<<2048, 256>>>(static_cast<const float4*>(src), static_cast<float4*>(dst), n16);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    return ret(e);
}

// n uniform doubles in (0, 1) into device out, stream `stream` of seed
int es_uniform(double* out, long long n, unsigned int seed, unsigned int stream)
{
    if (n <= 0)
        return 0;
    uniform_kernel<<<(unsigned)((n + 255) / 256), 256>>>(out, n, seed, stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    return ret(e);
}

}  // extern "C"
