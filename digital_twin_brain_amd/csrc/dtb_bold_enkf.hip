// dtb_bold_enkf.hip -- Balloon-Windkessel BOLD step and the diffusion EnKF
// analysis on device arrays.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "dtb_internal.hpp"

namespace dtb {

// models/bold_model_pytorch.py:69-77, in the reference's update order:
// d_s from the old s and f_in; q and v from the old f_in, q, v; f_in += s dt
// with the old s, clamped at 0; s += d_s dt; output from the new q, v.
template <typename T>
__global__ void bold_kernel(int64_t n, T eps, T ts, T tf, T t0, T div_alpha, T E0, T V0, T dt, T* __restrict__ s,
                            T* __restrict__ q, T* __restrict__ v, T* __restrict__ f, const T* __restrict__ u,
                            T* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T si = s[i], qi = q[i], vi = v[i], fi = f[i];
    const T d_s = eps * u[i] - si / ts - (fi - T(1)) / tf;
    const T q_part = fi > T(0) ? T(1) - pow(T(1) - E0, T(1) / fi) : T(1);
    const T qn = qi + (fi * q_part / E0 - qi * pow(vi, div_alpha - T(1))) / t0 * dt;
    const T vn = vi + (fi - pow(vi, div_alpha)) / t0 * dt;
    T fn = fi + si * dt;
    fn = fn > T(0) ? fn : T(0);
    const T sn = si + d_s * dt;
    s[i] = sn;
    q[i] = qn;
    v[i] = vn;
    f[i] = fn;
    out[i] = V0 * (T(7) * E0 * (T(1) - qn) + T(2) * (T(1) - qn / vn) + (T(2) * E0 - T(0.2)) * (T(1) - vn));
}

// `steps` consecutive steps of bold_kernel with a constant input per voxel,
// the state kept in registers: the same per-step arithmetic in the same order.
__global__ void bold_run_kernel(int64_t n, int32_t steps, double eps, double ts, double tf, double t0,
                                double div_alpha, double E0, double V0, double dt, double* __restrict__ s,
                                double* __restrict__ q, double* __restrict__ v, double* __restrict__ f,
                                const double* __restrict__ u, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double si = s[i], qi = q[i], vi = v[i], fi = f[i];
    const double ui = u[i];
    double o = out[i];
    for (int32_t k = 0; k < steps; ++k) {
        const double d_s = eps * ui - si / ts - (fi - 1.0) / tf;
        const double q_part = fi > 0.0 ? 1.0 - pow(1.0 - E0, 1.0 / fi) : 1.0;
        const double qn = qi + (fi * q_part / E0 - qi * pow(vi, div_alpha - 1.0)) / t0 * dt;
        const double vn = vi + (fi - pow(vi, div_alpha)) / t0 * dt;
        double fn = fi + si * dt;
        fn = fn > 0.0 ? fn : 0.0;
        si = si + d_s * dt;
        qi = qn;
        vi = vn;
        fi = fn;
        o = V0 * (7.0 * E0 * (1.0 - qn) + 2.0 * (1.0 - qn / vn) + (2.0 * E0 - 0.2) * (1.0 - vn));
    }
    s[i] = si;
    q[i] = qi;
    v[i] = vi;
    f[i] = fi;
    out[i] = o;
}

template <typename T>
static int bold_step(int64_t n, const double* par, T* s, T* q, T* v, T* f, const T* u, T* out, void* stream)
{
    DTB_REQUIRE(n >= 0, "n must be >= 0");
    DTB_REQUIRE(par != nullptr, "par is NULL");
    if (n == 0) return DTB_OK;
    DTB_REQUIRE(s && q && v && f && u && out, "state, input and output arrays are required");
    hipLaunchKernelGGL(bold_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                       (T)par[0], (T)par[1], (T)par[2], (T)par[3], (T)(1.0 / par[4]), (T)par[5], (T)par[6], (T)par[7],
                       s, q, v, f, u, out);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

// ------------------------------------------------------------------- EnKF
// data_assimilation/DataAssimilation.py:40-60.  w[E][V][S], last state =
// BOLD.  wd = w - mean_e(w); c_v = sum_e wd_ev^2/(E-1) + sigma;
// K_ev = wd_ev(last)/c_v/(E-1); innov_ev = y_v + sqrt(sigma) xi_ev - w_ev(last)
// w += rho innov_ev sum_e' K_e'v wd_e'vs + (1-rho) sum_e' A_ee' wd_e'vs,
// A_ee' = sum_v innov_ev K_e'v / V.
__global__ void enkf_mean_kernel(int32_t E, int32_t V, int32_t S, const double* __restrict__ w, double* __restrict__ mean)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; // (v, s)
    if (x >= (int64_t)V * S) return;
    double acc = 0.0;
    for (int32_t e = 0; e < E; ++e) acc += w[(int64_t)e * V * S + x];
    mean[x] = acc / E;
}

__global__ void enkf_diff_kernel(int32_t E, int32_t V, int32_t S, const double* __restrict__ w,
                                 const double* __restrict__ mean, double* __restrict__ wd)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)E * V * S) return;
    wd[x] = w[x] - mean[x % ((int64_t)V * S)];
}

__global__ void enkf_gain_kernel(int32_t E, int32_t V, int32_t S, const double* __restrict__ w,
                                 const double* __restrict__ wd, double sigma, const double* __restrict__ y,
                                 const double* __restrict__ xi, double* __restrict__ temp, double* __restrict__ innov)
{
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double c = 0.0;
    for (int32_t e = 0; e < E; ++e) {
        const double d = wd[((int64_t)e * V + v) * S + (S - 1)];
        c += d * d;
    }
    c = c / (E - 1) + sigma;
    const double sq = sqrt(sigma);
    for (int32_t e = 0; e < E; ++e) {
        const int64_t ev = (int64_t)e * V + v;
        temp[ev] = wd[ev * S + (S - 1)] / c / (E - 1);
        innov[ev] = y[v] + sq * xi[ev] - w[ev * S + (S - 1)];
    }
}

__global__ void enkf_local_kernel(int32_t E, int32_t V, int32_t S, const double* __restrict__ wd,
                                  const double* __restrict__ temp, double* __restrict__ gsum)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; // (v, s)
    if (x >= (int64_t)V * S) return;
    const int32_t v = (int32_t)(x / S);
    double acc = 0.0;
    for (int32_t e = 0; e < E; ++e) acc += temp[(int64_t)e * V + v] * wd[(int64_t)e * V * S + x];
    gsum[x] = acc;
}

__global__ void enkf_global_kernel(int32_t E, int32_t V, const double* __restrict__ innov,
                                   const double* __restrict__ temp, double* __restrict__ A)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; // (e, e')
    if (x >= (int64_t)E * E) return;
    const int32_t e = (int32_t)(x / E), f = (int32_t)(x % E);
    double acc = 0.0;
    for (int32_t v = 0; v < V; ++v) acc += innov[(int64_t)e * V + v] * temp[(int64_t)f * V + v];
    A[x] = acc / V;
}

__global__ void enkf_update_kernel(int32_t E, int32_t V, int32_t S, double rho, const double* __restrict__ wd,
                                   const double* __restrict__ innov, const double* __restrict__ gsum,
                                   const double* __restrict__ A, double* __restrict__ w)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; // (e, v, s)
    if (x >= (int64_t)E * V * S) return;
    const int64_t vs = x % ((int64_t)V * S);
    const int32_t e = (int32_t)(x / ((int64_t)V * S));
    const int32_t v = (int32_t)(vs / S);
    double acc = 0.0;
    for (int32_t f = 0; f < E; ++f) acc += A[(int64_t)e * E + f] * wd[(int64_t)f * V * S + vs];
    w[x] += rho * innov[(int64_t)e * V + v] * gsum[vs] + (1.0 - rho) * acc;
}

// The six analysis kernels on a caller-chosen scratch of enkf_scratch_doubles
// doubles: mean [V*S], wd [E*V*S], temp [E*V], innov [E*V], gsum [V*S],
// A [E*E].  Every scratch element is written before it is read.
static size_t enkf_scratch_doubles(int32_t E, int32_t V, int32_t S)
{
    const size_t EVS = (size_t)E * V * S, VS = (size_t)V * S, EV = (size_t)E * V, EE = (size_t)E * E;
    return VS + EVS + 2 * EV + VS + EE;
}

static int enkf_launch(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y, double rho,
                       const double* xi, hipStream_t st, double* scratch)
{
    const size_t EVS = (size_t)E * V * S, VS = (size_t)V * S, EV = (size_t)E * V, EE = (size_t)E * E;
    double* mean = scratch;
    double* wd = mean + VS;
    double* temp = wd + EVS;
    double* innov = temp + EV;
    double* gsum = innov + EV;
    double* A = gsum + VS;
    auto g = [](size_t n) { return dim3((unsigned)std::max<size_t>(1, (n + 255) / 256)); };
    hipLaunchKernelGGL(enkf_mean_kernel, g(VS), dim3(256), 0, st, E, V, S, w, mean);
    hipLaunchKernelGGL(enkf_diff_kernel, g(EVS), dim3(256), 0, st, E, V, S, w, mean, wd);
    hipLaunchKernelGGL(enkf_gain_kernel, g((size_t)V), dim3(256), 0, st, E, V, S, w, wd, sigma, y, xi, temp, innov);
    hipLaunchKernelGGL(enkf_local_kernel, g(VS), dim3(256), 0, st, E, V, S, wd, temp, gsum);
    hipLaunchKernelGGL(enkf_global_kernel, g(EE), dim3(256), 0, st, E, V, innov, temp, A);
    hipLaunchKernelGGL(enkf_update_kernel, g(EVS), dim3(256), 0, st, E, V, S, rho, wd, innov, gsum, A, w);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

// Scratch for the analysis, kept per device and grown on demand.
struct EnkfScratch {
    double* p = nullptr;
    size_t cap = 0;
};
static std::mutex g_enkf_mu;
static EnkfScratch g_enkf[64];

static int enkf(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y, double rho,
                const double* xi, void* stream)
{
    DTB_REQUIRE(E >= 2 && V >= 1 && S >= 1, "EnKF needs E >= 2, V >= 1, S >= 1");
    DTB_REQUIRE(w && y && xi, "w, y and xi are required");
    int dev = 0;
    DTB_TRY_HIP(hipGetDevice(&dev));
    DTB_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
    const size_t need = enkf_scratch_doubles(E, V, S);
    std::lock_guard<std::mutex> lk(g_enkf_mu);
    EnkfScratch& sc = g_enkf[dev];
    if (sc.cap < need) {
        if (sc.p) {
            DTB_TRY_HIP(hipDeviceSynchronize());
            DTB_TRY_HIP(hipFree(sc.p));
        }
        sc.p = nullptr;
        sc.cap = 0;
        DTB_TRY_HIP(hipMalloc(&sc.p, need * sizeof(double)));
        sc.cap = need;
    }
    hipStream_t st = (hipStream_t)stream;
    DTB_TRY(enkf_launch(E, V, S, w, sigma, y, rho, xi, st, sc.p));
    // the scratch is shared by later calls on this device
    DTB_TRY_HIP(hipStreamSynchronize(st));
    return DTB_OK;
}

// Allocation-free analysis: the scratch is a fixed device array, one copy per
// GPU, held until the process ends.  Its host-side name is not a device
// address, so each device's address comes from hipGetSymbolAddress, taken
// once per device with that device current, and only that address is handed
// to the kernels.
constexpr size_t kEnkfScratchMax = (size_t)1 << 19; // doubles (4 MiB): E = 256, V = 90, S = 10 needs 343,816
static __device__ __align__(256) double scratch[kEnkfScratchMax];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {nullptr};

static int enkf_noalloc(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y, double rho,
                        const double* xi, void* stream)
{
    DTB_REQUIRE(E >= 2 && V >= 1 && S >= 1, "EnKF needs E >= 2, V >= 1, S >= 1");
    DTB_REQUIRE(w && y && xi, "w, y and xi are required");
    DTB_REQUIRE((size_t)E <= kEnkfScratchMax && (size_t)V <= kEnkfScratchMax && (size_t)S <= kEnkfScratchMax,
                "allocation-free EnKF: E=%d, V=%d, S=%d exceeds the fixed scratch of %zu doubles "
                "(use dtb_enkf_analysis_f64)", E, V, S, kEnkfScratchMax);
    const size_t need = enkf_scratch_doubles(E, V, S);
    DTB_REQUIRE(need <= kEnkfScratchMax,
                "allocation-free EnKF: E=%d, V=%d, S=%d needs %zu scratch doubles, the fixed scratch holds %zu "
                "(use dtb_enkf_analysis_f64)", E, V, S, need, kEnkfScratchMax);
    int dev = 0;
    DTB_TRY_HIP(hipGetDevice(&dev));
    DTB_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
    double* p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* a = nullptr;
            const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !a) {
                set_error("hipGetSymbolAddress of the EnKF scratch failed on device %d: %s", dev, hipGetErrorString(e));
                return DTB_RUNTIME_ERROR;
            }
            g_scratch_ptr[dev] = static_cast<double*>(a);
        }
        p = g_scratch_ptr[dev];
    }
    return enkf_launch(E, V, S, w, sigma, y, rho, xi, (hipStream_t)stream, p);
}

} // namespace dtb

extern "C" {

int dtb_bold_step_f32(int64_t n, const double* par, float* s, float* q, float* v, float* f, const float* u,
                      float* out, void* stream)
{
    return dtb::bold_step<float>(n, par, s, q, v, f, u, out, stream);
}

int dtb_bold_step_f64(int64_t n, const double* par, double* s, double* q, double* v, double* f, const double* u,
                      double* out, void* stream)
{
    return dtb::bold_step<double>(n, par, s, q, v, f, u, out, stream);
}

int dtb_bold_run_f64(int64_t n, int32_t steps, const double* par, double* s, double* q, double* v, double* f,
                     const double* u, double* out, void* stream)
{
    DTB_REQUIRE(n >= 0 && steps >= 0, "n and steps must be >= 0");
    DTB_REQUIRE(par != nullptr, "par is NULL");
    if (n == 0 || steps == 0) return DTB_OK;
    DTB_REQUIRE(s && q && v && f && u && out, "state, input and output arrays are required");
    hipLaunchKernelGGL(dtb::bold_run_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                       steps, par[0], par[1], par[2], par[3], 1.0 / par[4], par[5], par[6], par[7], s, q, v, f, u, out);
    DTB_TRY_HIP(hipGetLastError());
    return DTB_OK;
}

int dtb_enkf_analysis_f64(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y, double rho,
                          const double* xi, void* stream)
{
    return dtb::enkf(E, V, S, w, sigma, y, rho, xi, stream);
}

int dtb_enkf_analysis_f64_noalloc(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y,
                                  double rho, const double* xi, void* stream)
{
    return dtb::enkf_noalloc(E, V, S, w, sigma, y, rho, xi, stream);
}

} // extern "C"
