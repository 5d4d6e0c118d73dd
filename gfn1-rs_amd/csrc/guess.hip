// Learned SCC start charges: descriptor, per-atom MLP inference and training
// kernels (FP32), Adam, and the host-side model.
#include "guess.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

namespace gfn1 {
namespace {

constexpr double kBohrToAngstrom = 1.0 / 1.8897261246257702;
constexpr double kCutoff = 5.0;                         // A
constexpr double kMu0 = 0.5, kMuStep = 0.3;              // A
constexpr double kEta = 1.0 / (2.0 * kMuStep * kMuStep);  // A^-2
constexpr int kRadial = 16;
constexpr int kAtomsPerGroup = 64;
// training work-group LDS rows (odd strides: no bank conflicts)
constexpr int kXs = kGuessFeat + 1;
constexpr int kHs = kGuessHidden + 1;

#define GHIP(expr)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return Status::error(kErrDevice, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

__device__ __forceinline__ int slot_of(int z) { return z == 1 ? 0 : z == 6 ? 1 : z == 7 ? 2 : z == 8 ? 3 : -1; }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// One thread per atom; neighbours are the other atoms of the same instance.
__global__ void k_guess_features(int natoms, const int* __restrict__ z, const double* __restrict__ x,
                                 const int* __restrict__ inst, const int* __restrict__ inst_atom0,
                                 float* __restrict__ feat) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    float* f = feat + (size_t)a * kGuessFeat;
    const int sa = slot_of(z[a]);
    for (int e = 0; e < 4; ++e) f[e] = sa == e ? 1.0f : 0.0f;
    const int i = inst[a];
    const int j0 = inst_atom0[i], j1 = inst_atom0[i + 1];
    const double xa = x[3 * a], ya = x[3 * a + 1], za = x[3 * a + 2];
    for (int e = 0; e < 4; ++e) {
        double acc[kRadial];
#pragma unroll
        for (int k = 0; k < kRadial; ++k) acc[k] = 0.0;
        for (int j = j0; j < j1; ++j) {
            if (j == a || slot_of(z[j]) != e) continue;
            const double dx = x[3 * j] - xa, dy = x[3 * j + 1] - ya, dz = x[3 * j + 2] - za;
            const double r = sqrt(dx * dx + dy * dy + dz * dz) * kBohrToAngstrom;
            if (r >= kCutoff) continue;
            const double fc = 0.5 * (cos(M_PI * r / kCutoff) + 1.0);
#pragma unroll
            for (int k = 0; k < kRadial; ++k) {
                const double d = r - (kMu0 + kMuStep * k);
                acc[k] += exp(-kEta * d * d) * fc;
            }
        }
#pragma unroll
        for (int k = 0; k < kRadial; ++k) f[4 + kRadial * e + k] = (float)acc[k];
    }
}

// Inference: one thread per atom, parameters in LDS, each thread's hidden
// layer in its own LDS row.
__global__ void __launch_bounds__(kAtomsPerGroup) k_guess_forward(int natoms, const float* __restrict__ theta,
                                                                  const float* __restrict__ norm,
                                                                  const float* __restrict__ feat,
                                                                  float* __restrict__ out) {
    __shared__ float th[kGuessParams];
    __shared__ float nm[2 * kGuessFeat];
    __shared__ float H1[kAtomsPerGroup * kHs];
    for (int p = threadIdx.x; p < kGuessParams; p += blockDim.x) th[p] = theta[p];
    for (int p = threadIdx.x; p < 2 * kGuessFeat; p += blockDim.x) nm[p] = norm[p];
    __syncthreads();
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    float* h1 = H1 + threadIdx.x * kHs;
    float xv[kGuessFeat];
#pragma unroll
    for (int i = 0; i < kGuessFeat; ++i) xv[i] = (feat[(size_t)a * kGuessFeat + i] - nm[i]) * nm[kGuessFeat + i];
    for (int j = 0; j < kGuessHidden; ++j) {
        float s = th[kGoB1 + j];
#pragma unroll
        for (int i = 0; i < kGuessFeat; ++i) s += th[kGoW1 + j * kGuessFeat + i] * xv[i];
        h1[j] = silu(s);
    }
    float y0 = th[kGoB3], y1 = th[kGoB3 + 1];
    for (int j = 0; j < kGuessHidden; ++j) {
        float s = th[kGoB2 + j];
#pragma unroll 16
        for (int i = 0; i < kGuessHidden; ++i) s += th[kGoW2 + j * kGuessHidden + i] * h1[i];
        const float h2 = silu(s);
        y0 += th[kGoW3 + j] * h2;
        y1 += th[kGoW3 + kGuessHidden + j] * h2;
    }
    out[2 * (size_t)a] = y0;
    out[2 * (size_t)a + 1] = y1;
}

// Training: forward + backward for up to 64 atoms (idx[blockIdx.x*64 + lane]),
// loss 1/(2 nb) sum (y - t)^2; every thread keeps its atom's activations in its
// own LDS rows; the work-group's parameter-gradient sums go to
// partial[blockIdx.x] (deterministic; reduced by k_guess_reduce).
__global__ void __launch_bounds__(kAtomsPerGroup) k_guess_backward(int nb, const int* __restrict__ idx,
                                                                   const float* __restrict__ theta,
                                                                   const float* __restrict__ norm,
                                                                   const float* __restrict__ feat,
                                                                   const float* __restrict__ labels,
                                                                   float* __restrict__ partial,
                                                                   float* __restrict__ loss_partial) {
    extern __shared__ float sm[];
    float* th = sm;                                      // kGuessParams
    float* nm = th + kGuessParams;                       // 2 * kGuessFeat
    float* X = nm + 2 * kGuessFeat;                      // 64 x kXs  standardised inputs
    float* H1 = X + kAtomsPerGroup * kXs;                // 64 x kHs  layer-1 outputs
    float* H2 = H1 + kAtomsPerGroup * kHs;               //           layer-2 outputs
    float* DZ1 = H2 + kAtomsPerGroup * kHs;              //           SiLU'(z1), then dL/dz1
    float* DZ2 = DZ1 + kAtomsPerGroup * kHs;             //           SiLU'(z2), then dL/dz2
    float* DY = DZ2 + kAtomsPerGroup * kHs;              // 64 x 2    dL/dy
    const int tid = threadIdx.x;
    for (int p = tid; p < kGuessParams; p += blockDim.x) th[p] = theta[p];
    for (int p = tid; p < 2 * kGuessFeat; p += blockDim.x) nm[p] = norm[p];
    __syncthreads();
    const int slot = blockIdx.x * kAtomsPerGroup + tid;
    float* xr = X + tid * kXs;
    float* h1 = H1 + tid * kHs;
    float* h2 = H2 + tid * kHs;
    float* d1 = DZ1 + tid * kHs;
    float* d2 = DZ2 + tid * kHs;
    float loss = 0.0f;
    if (slot < nb) {
        const int a = idx[slot];
        float xv[kGuessFeat];
#pragma unroll
        for (int i = 0; i < kGuessFeat; ++i) {
            xv[i] = (feat[(size_t)a * kGuessFeat + i] - nm[i]) * nm[kGuessFeat + i];
            xr[i] = xv[i];
        }
        for (int j = 0; j < kGuessHidden; ++j) {
            float s = th[kGoB1 + j];
#pragma unroll
            for (int i = 0; i < kGuessFeat; ++i) s += th[kGoW1 + j * kGuessFeat + i] * xv[i];
            h1[j] = silu(s);
            d1[j] = dsilu(s);
        }
        float y0 = th[kGoB3], y1 = th[kGoB3 + 1];
        for (int j = 0; j < kGuessHidden; ++j) {
            float s = th[kGoB2 + j];
#pragma unroll 16
            for (int i = 0; i < kGuessHidden; ++i) s += th[kGoW2 + j * kGuessHidden + i] * h1[i];
            const float hv = silu(s);
            h2[j] = hv;
            d2[j] = dsilu(s);
            y0 += th[kGoW3 + j] * hv;
            y1 += th[kGoW3 + kGuessHidden + j] * hv;
        }
        const float r0 = y0 - labels[2 * (size_t)a], r1 = y1 - labels[2 * (size_t)a + 1];
        loss = 0.5f * (r0 * r0 + r1 * r1) / (float)nb;
        const float e0 = r0 / (float)nb, e1 = r1 / (float)nb;
        DY[2 * tid] = e0;
        DY[2 * tid + 1] = e1;
        for (int j = 0; j < kGuessHidden; ++j)
            d2[j] *= th[kGoW3 + j] * e0 + th[kGoW3 + kGuessHidden + j] * e1;
        for (int i = 0; i < kGuessHidden; ++i) {
            float s = 0.0f;
#pragma unroll 16
            for (int j = 0; j < kGuessHidden; ++j) s += th[kGoW2 + j * kGuessHidden + i] * d2[j];
            d1[i] *= s;
        }
    } else {
        for (int i = 0; i < kGuessFeat; ++i) xr[i] = 0.0f;
        for (int j = 0; j < kGuessHidden; ++j) h1[j] = h2[j] = d1[j] = d2[j] = 0.0f;
        DY[2 * tid] = DY[2 * tid + 1] = 0.0f;
    }
    __syncthreads();
    float* gp = partial + (size_t)blockIdx.x * kGuessParams;
    for (int p = tid; p < kGuessParams; p += blockDim.x) {
        float s = 0.0f;
        if (p < kGoB1) {
            const int j = p / kGuessFeat, i = p % kGuessFeat;
            for (int a = 0; a < kAtomsPerGroup; ++a) s += DZ1[a * kHs + j] * X[a * kXs + i];
        } else if (p < kGoW2) {
            const int j = p - kGoB1;
            for (int a = 0; a < kAtomsPerGroup; ++a) s += DZ1[a * kHs + j];
        } else if (p < kGoB2) {
            const int j = (p - kGoW2) / kGuessHidden, i = (p - kGoW2) % kGuessHidden;
            for (int a = 0; a < kAtomsPerGroup; ++a) s += DZ2[a * kHs + j] * H1[a * kHs + i];
        } else if (p < kGoW3) {
            const int j = p - kGoB2;
            for (int a = 0; a < kAtomsPerGroup; ++a) s += DZ2[a * kHs + j];
        } else if (p < kGoB3) {
            const int o = (p - kGoW3) / kGuessHidden, i = (p - kGoW3) % kGuessHidden;
            for (int a = 0; a < kAtomsPerGroup; ++a) s += DY[2 * a + o] * H2[a * kHs + i];
        } else {
            const int o = p - kGoB3;
            for (int a = 0; a < kAtomsPerGroup; ++a) s += DY[2 * a + o];
        }
        gp[p] = s;
    }
    // work-group loss (one wavefront)
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (tid == 0) loss_partial[blockIdx.x] = loss;
}

// grad = sum of the work-group partials (fixed order); loss_acc += weight * sum(loss partials)
__global__ void k_guess_reduce(int nwg, const float* __restrict__ partial, float* __restrict__ grad,
                               const float* __restrict__ loss_partial, double* loss_acc, double weight) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kGuessParams) {
        float s = 0.0f;
        for (int w = 0; w < nwg; ++w) s += partial[(size_t)w * kGuessParams + p];
        grad[p] = s;
    }
    if (p == 0 && loss_acc) {
        double l = 0.0;
        for (int w = 0; w < nwg; ++w) l += loss_partial[w];
        *loss_acc += weight * l;
    }
}

__global__ void k_guess_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ g, float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGuessParams) return;
    constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    theta[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// Start vector from the per-atom predictions: shells of atom a take pred[2a],
// pred[2a+1]; the instance is shifted uniformly so that its shell charges sum
// to its charge.  Instances the model does not cover keep q.
__global__ void k_guess_apply(int ninst, const int* __restrict__ usable, const int* __restrict__ atom0,
                              const int* __restrict__ atom_sh0, const double* __restrict__ charge,
                              const float* __restrict__ pred, double* q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst || !usable[i]) return;
    const int a0 = atom0[i], a1 = atom0[i + 1];
    double sum = 0.0;
    for (int a = a0; a < a1; ++a) sum += (double)pred[2 * a] + (double)pred[2 * a + 1];
    const int nsh = 2 * (a1 - a0);  // the model covers instances whose atoms all have two shells
    const double shift = nsh > 0 ? (sum - charge[i]) / nsh : 0.0;
    for (int a = a0; a < a1; ++a) {
        q[atom_sh0[a]] = (double)pred[2 * a] - shift;
        q[atom_sh0[a] + 1] = (double)pred[2 * a + 1] - shift;
    }
}

size_t backward_lds_bytes() {
    return ((size_t)kGuessParams + 2 * kGuessFeat + kAtomsPerGroup * kXs + 4 * kAtomsPerGroup * kHs +
            2 * kAtomsPerGroup) *
           sizeof(float);
}

hipError_t set_backward_lds() {
    static bool done = false;
    if (done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void*)k_guess_backward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)backward_lds_bytes());
    if (e == hipSuccess) done = true;
    return e;
}

// SplitMix64 stream for initialisation and shuffling.
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Device copies of a GuessAtoms set.
struct DeviceAtoms {
    int natoms = 0;
    int *z = nullptr, *inst = nullptr, *inst_atom0 = nullptr;
    double* x = nullptr;
    float* feat = nullptr;
    ~DeviceAtoms() {
        (void)hipFree(z);
        (void)hipFree(inst);
        (void)hipFree(inst_atom0);
        (void)hipFree(x);
        (void)hipFree(feat);
    }
    Status upload(const GuessAtoms& a) {
        natoms = (int)a.z.size();
        if (natoms == 0) return Status{};
        GHIP(hipMalloc(&z, natoms * sizeof(int)));
        GHIP(hipMalloc(&inst, natoms * sizeof(int)));
        GHIP(hipMalloc(&inst_atom0, a.inst_atom0.size() * sizeof(int)));
        GHIP(hipMalloc(&x, 3 * (size_t)natoms * sizeof(double)));
        GHIP(hipMalloc(&feat, (size_t)natoms * kGuessFeat * sizeof(float)));
        GHIP(hipMemcpy(z, a.z.data(), natoms * sizeof(int), hipMemcpyHostToDevice));
        GHIP(hipMemcpy(inst, a.inst.data(), natoms * sizeof(int), hipMemcpyHostToDevice));
        GHIP(hipMemcpy(inst_atom0, a.inst_atom0.data(), a.inst_atom0.size() * sizeof(int), hipMemcpyHostToDevice));
        GHIP(hipMemcpy(x, a.x.data(), 3 * (size_t)natoms * sizeof(double), hipMemcpyHostToDevice));
        GHIP(guess_launch_features(0, natoms, z, x, inst, inst_atom0, feat));
        return Status{};
    }
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

std::vector<int> usable_atoms(const GuessAtoms& a) {
    std::vector<int> out;
    for (int i = 0; i < (int)a.z.size(); ++i)
        if (a.usable[a.inst[i]]) out.push_back(i);
    return out;
}

}  // namespace

hipError_t guess_launch_features(hipStream_t s, int natoms, const int* z, const double* x, const int* inst,
                                 const int* inst_atom0, float* feat) {
    if (natoms <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guess_features, dim3((natoms + 127) / 128), dim3(128), 0, s, natoms, z, x, inst,
                       inst_atom0, feat);
    return hipGetLastError();
}

hipError_t guess_launch_forward(hipStream_t s, int natoms, const float* theta, const float* norm, const float* feat,
                                float* out) {
    if (natoms <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guess_forward, dim3((natoms + kAtomsPerGroup - 1) / kAtomsPerGroup), dim3(kAtomsPerGroup),
                       0, s, natoms, theta, norm, feat, out);
    return hipGetLastError();
}

hipError_t guess_launch_apply(hipStream_t s, int ninst, const int* usable, const int* atom0, const int* atom_sh0,
                              const double* charge, const float* pred, double* q) {
    if (ninst <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guess_apply, dim3((ninst + 127) / 128), dim3(128), 0, s, ninst, usable, atom0, atom_sh0,
                       charge, pred, q);
    return hipGetLastError();
}

Status guess_features(const GuessAtoms& atoms, std::vector<float>& out) {
    DeviceAtoms da;
    Status st = da.upload(atoms);
    if (!st.ok()) return st;
    out.assign((size_t)da.natoms * kGuessFeat, 0.0f);
    if (da.natoms) GHIP(hipMemcpy(out.data(), da.feat, out.size() * sizeof(float), hipMemcpyDeviceToHost));
    return Status{};
}

GuessModel::~GuessModel() {
    (void)hipFree(d_theta_);
    (void)hipFree(d_norm_);
}

Status GuessModel::init_random(uint64_t seed) {
    SplitMix rng{seed};
    theta_.assign(kGuessParams, 0.0f);
    auto fill = [&](int off, int rows, int cols, double scale) {
        for (int k = 0; k < rows * cols; ++k) theta_[off + k] = (float)(rng.normal() * scale / std::sqrt((double)cols));
    };
    fill(kGoW1, kGuessHidden, kGuessFeat, 1.0);
    fill(kGoW2, kGuessHidden, kGuessHidden, 1.0);
    fill(kGoW3, kGuessOut, kGuessHidden, 0.1);
    norm_.assign(2 * kGuessFeat, 0.0f);
    for (int i = 0; i < kGuessFeat; ++i) norm_[kGuessFeat + i] = 1.0f;
    dirty_ = true;
    return upload_if_device();
}

// Parameters go to the current device as soon as the model exists when a
// device is present, so that later calls (the no-allocation prediction among
// them) find them resident; without a device they are uploaded on first use.
Status GuessModel::upload_if_device() const {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return Status{};
    }
    return upload();
}

Status GuessModel::set_params(const float* theta) {
    theta_.assign(theta, theta + kGuessParams);
    if (norm_.size() != 2 * kGuessFeat) {
        norm_.assign(2 * kGuessFeat, 0.0f);
        for (int i = 0; i < kGuessFeat; ++i) norm_[kGuessFeat + i] = 1.0f;
    }
    dirty_ = true;
    return Status{};
}

Status GuessModel::upload() const {
    if (!dirty_ && d_theta_) return Status{};
    if (!d_theta_) GHIP(hipMalloc(&d_theta_, kGuessParams * sizeof(float)));
    if (!d_norm_) GHIP(hipMalloc(&d_norm_, 2 * kGuessFeat * sizeof(float)));
    GHIP(hipMemcpy(d_theta_, theta_.data(), kGuessParams * sizeof(float), hipMemcpyHostToDevice));
    GHIP(hipMemcpy(d_norm_, norm_.data(), 2 * kGuessFeat * sizeof(float), hipMemcpyHostToDevice));
    dirty_ = false;
    return Status{};
}

Status GuessModel::device_ptrs(const float** theta, const float** norm) const {
    if (theta_.size() != (size_t)kGuessParams) return Status::error(kErrInvalidInput, "guess model has no parameters");
    Status st = upload();
    if (!st.ok()) return st;
    *theta = d_theta_;
    *norm = d_norm_;
    return Status{};
}

// File: 8-byte magic "GFN1MLP1", int32 parameter count, int32 feature count,
// float32 parameters, float32 mean[68], float32 inverse spread[68].
Status GuessModel::save(const std::string& path) const {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return Status::error(kErrIo, "cannot write " + path);
    const int32_t hdr[2] = {kGuessParams, kGuessFeat};
    bool ok = std::fwrite("GFN1MLP1", 1, 8, f) == 8 && std::fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
              std::fwrite(theta_.data(), sizeof(float), kGuessParams, f) == (size_t)kGuessParams &&
              std::fwrite(norm_.data(), sizeof(float), 2 * kGuessFeat, f) == (size_t)(2 * kGuessFeat);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? Status{} : Status::error(kErrIo, "short write to " + path);
}

Status GuessModel::load(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return Status::error(kErrIo, "cannot read " + path);
    char magic[8];
    int32_t hdr[2];
    std::vector<float> th(kGuessParams), nm(2 * kGuessFeat);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "GFN1MLP1", 8) == 0 &&
              std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == kGuessParams && hdr[1] == kGuessFeat &&
              std::fread(th.data(), sizeof(float), kGuessParams, f) == (size_t)kGuessParams &&
              std::fread(nm.data(), sizeof(float), 2 * kGuessFeat, f) == (size_t)(2 * kGuessFeat);
    std::fclose(f);
    if (!ok) return Status::error(kErrParse, "not a guess model file: " + path);
    theta_ = th;
    norm_ = nm;
    dirty_ = true;
    return upload_if_device();
}

Status GuessModel::predict_atoms(const GuessAtoms& atoms, std::vector<float>& out) const {
    const float *th, *nm;
    Status st = device_ptrs(&th, &nm);
    if (!st.ok()) return st;
    DeviceAtoms da;
    st = da.upload(atoms);
    if (!st.ok()) return st;
    out.assign(2 * (size_t)da.natoms, 0.0f);
    if (!da.natoms) return Status{};
    DevBuf<float> y;
    GHIP(y.alloc(2 * (size_t)da.natoms));
    GHIP(guess_launch_forward(0, da.natoms, th, nm, da.feat, y.p));
    GHIP(hipMemcpy(out.data(), y.p, out.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < da.natoms; ++i)
        if (!atoms.usable[atoms.inst[i]]) out[2 * i] = out[2 * i + 1] = 0.0f;
    return Status{};
}

Status GuessModel::loss_gradient(const GuessAtoms& atoms, const std::vector<float>& labels, double* loss,
                                 std::vector<float>& grad) const {
    const float *th, *nm;
    Status st = device_ptrs(&th, &nm);
    if (!st.ok()) return st;
    DeviceAtoms da;
    st = da.upload(atoms);
    if (!st.ok()) return st;
    const std::vector<int> use = usable_atoms(atoms);
    const int nb = (int)use.size();
    if (nb == 0) return Status::error(kErrInvalidInput, "no atoms the guess model covers (H, C, N, O)");
    const int nwg = (nb + kAtomsPerGroup - 1) / kAtomsPerGroup;
    DevBuf<int> idx;
    DevBuf<float> lab, part, lpart, g;
    DevBuf<double> lacc;
    GHIP(idx.alloc(nb));
    GHIP(lab.alloc(labels.size()));
    GHIP(part.alloc((size_t)nwg * kGuessParams));
    GHIP(lpart.alloc(nwg));
    GHIP(g.alloc(kGuessParams));
    GHIP(lacc.alloc(1));
    GHIP(hipMemcpy(idx.p, use.data(), nb * sizeof(int), hipMemcpyHostToDevice));
    GHIP(hipMemcpy(lab.p, labels.data(), labels.size() * sizeof(float), hipMemcpyHostToDevice));
    GHIP(hipMemset(lacc.p, 0, sizeof(double)));
    GHIP(set_backward_lds());
    hipLaunchKernelGGL(k_guess_backward, dim3(nwg), dim3(kAtomsPerGroup), backward_lds_bytes(), 0, nb, idx.p, th, nm,
                       da.feat, lab.p, part.p, lpart.p);
    GHIP(hipGetLastError());
    hipLaunchKernelGGL(k_guess_reduce, dim3((kGuessParams + 255) / 256), dim3(256), 0, 0, nwg, part.p, g.p, lpart.p,
                       lacc.p, 1.0);
    GHIP(hipGetLastError());
    grad.assign(kGuessParams, 0.0f);
    GHIP(hipMemcpy(grad.data(), g.p, kGuessParams * sizeof(float), hipMemcpyDeviceToHost));
    GHIP(hipMemcpy(loss, lacc.p, sizeof(double), hipMemcpyDeviceToHost));
    return Status{};
}

Status GuessModel::train(const GuessAtoms& atoms, const std::vector<float>& labels, const GuessTrainOptions& opt,
                         std::vector<float>& loss_per_epoch) {
    if (theta_.size() != (size_t)kGuessParams) init_random(opt.seed);
    if (opt.epochs < 0 || opt.batch_atoms < 1 || !(opt.learning_rate > 0.0f))
        return Status::error(kErrInvalidInput, "invalid training options");
    DeviceAtoms da;
    Status st = da.upload(atoms);
    if (!st.ok()) return st;
    std::vector<int> use = usable_atoms(atoms);
    const int n = (int)use.size();
    if (n == 0) return Status::error(kErrInvalidInput, "no atoms the guess model covers (H, C, N, O)");
    // standardisation of the descriptor from the training atoms
    {
        std::vector<float> f((size_t)da.natoms * kGuessFeat);
        GHIP(hipMemcpy(f.data(), da.feat, f.size() * sizeof(float), hipMemcpyDeviceToHost));
        std::vector<double> mean(kGuessFeat, 0.0), var(kGuessFeat, 0.0);
        for (int a : use)
            for (int i = 0; i < kGuessFeat; ++i) mean[i] += f[(size_t)a * kGuessFeat + i];
        for (double& m : mean) m /= n;
        for (int a : use)
            for (int i = 0; i < kGuessFeat; ++i) {
                const double d = f[(size_t)a * kGuessFeat + i] - mean[i];
                var[i] += d * d;
            }
        norm_.assign(2 * kGuessFeat, 0.0f);
        for (int i = 0; i < kGuessFeat; ++i) {
            norm_[i] = (float)mean[i];
            norm_[kGuessFeat + i] = (float)(1.0 / (std::sqrt(var[i] / n) + 1e-3));
        }
        dirty_ = true;
    }
    const float *th_c, *nm;
    st = device_ptrs(&th_c, &nm);
    if (!st.ok()) return st;
    float* th = d_theta_;
    const int B = std::min(opt.batch_atoms, n);
    const int nwg_max = (B + kAtomsPerGroup - 1) / kAtomsPerGroup;
    DevBuf<int> idx;
    DevBuf<float> lab, part, lpart, g, m, v;
    DevBuf<double> lacc;
    GHIP(idx.alloc(n));
    GHIP(lab.alloc(labels.size()));
    GHIP(part.alloc((size_t)nwg_max * kGuessParams));
    GHIP(lpart.alloc(nwg_max));
    GHIP(g.alloc(kGuessParams));
    GHIP(m.alloc(kGuessParams));
    GHIP(v.alloc(kGuessParams));
    GHIP(lacc.alloc(std::max(opt.epochs, 1)));
    GHIP(hipMemcpy(lab.p, labels.data(), labels.size() * sizeof(float), hipMemcpyHostToDevice));
    GHIP(hipMemset(m.p, 0, kGuessParams * sizeof(float)));
    GHIP(hipMemset(v.p, 0, kGuessParams * sizeof(float)));
    GHIP(hipMemset(lacc.p, 0, std::max(opt.epochs, 1) * sizeof(double)));
    GHIP(set_backward_lds());
    SplitMix rng{opt.seed ^ 0x5eedULL};
    int step = 0;
    for (int ep = 0; ep < opt.epochs; ++ep) {
        for (int i = n - 1; i > 0; --i) std::swap(use[i], use[(int)(rng.next() % (uint64_t)(i + 1))]);
        GHIP(hipMemcpy(idx.p, use.data(), n * sizeof(int), hipMemcpyHostToDevice));
        for (int off = 0; off < n; off += B) {
            const int nb = std::min(B, n - off);
            const int nwg = (nb + kAtomsPerGroup - 1) / kAtomsPerGroup;
            hipLaunchKernelGGL(k_guess_backward, dim3(nwg), dim3(kAtomsPerGroup), backward_lds_bytes(), 0, nb,
                               idx.p + off, (const float*)th, nm, (const float*)da.feat, (const float*)lab.p, part.p,
                               lpart.p);
            GHIP(hipGetLastError());
            hipLaunchKernelGGL(k_guess_reduce, dim3((kGuessParams + 255) / 256), dim3(256), 0, 0, nwg,
                               (const float*)part.p, g.p, (const float*)lpart.p, lacc.p + ep, (double)nb / n);
            GHIP(hipGetLastError());
            ++step;
            const float c1 = 1.0f - std::pow(0.9f, (float)step), c2 = 1.0f - std::pow(0.999f, (float)step);
            hipLaunchKernelGGL(k_guess_adam, dim3((kGuessParams + 255) / 256), dim3(256), 0, 0, th, m.p, v.p,
                               (const float*)g.p, opt.learning_rate, c1, c2);
            GHIP(hipGetLastError());
        }
    }
    GHIP(hipDeviceSynchronize());
    std::vector<double> l(std::max(opt.epochs, 1));
    GHIP(hipMemcpy(l.data(), lacc.p, l.size() * sizeof(double), hipMemcpyDeviceToHost));
    loss_per_epoch.assign(opt.epochs, 0.0f);
    for (int ep = 0; ep < opt.epochs; ++ep) loss_per_epoch[ep] = (float)l[ep];
    GHIP(hipMemcpy(theta_.data(), th, kGuessParams * sizeof(float), hipMemcpyDeviceToHost));
    return Status{};
}

// ---------------------------------------------------------------------------
// Prediction without device allocation.  Its device memory is the file-scope
// array `scratch`: staged atom tables and positions, then the descriptors and
// network outputs, each region on a 256-byte boundary.  Every kernel writes
// all of its outputs, so nothing needs clearing between calls.  In host code
// `scratch` names a host-side stand-in; the device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
namespace {
constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaZ = 0;
constexpr size_t kNaInst = kNaZ + round256(sizeof(int) * kNoAllocMaxAtoms);
constexpr size_t kNaAtom0 = kNaInst + round256(sizeof(int) * kNoAllocMaxAtoms);
constexpr size_t kNaX = kNaAtom0 + round256(sizeof(int) * (kNoAllocMaxSystems + 1));
constexpr size_t kNaFeat = kNaX + round256(sizeof(double) * 3 * kNoAllocMaxAtoms);
constexpr size_t kNaOut = kNaFeat + round256(sizeof(float) * kGuessFeat * kNoAllocMaxAtoms);
constexpr size_t kNaBytes = kNaOut + round256(sizeof(float) * kGuessOut * kNoAllocMaxAtoms);

static __device__ __align__(256) unsigned char scratch[kNaBytes];

std::mutex g_scratch_mu;                  // guards the pointer table and serialises the calls
std::vector<unsigned char*> g_scratch_ptr;  // device address of `scratch`, per device id

// Device address of `scratch` on the current device (looked up on the first
// call on that device).  The caller holds g_scratch_mu.
Status scratch_address(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return Status::error(kErrDevice, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return Status::error(kErrDevice, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return Status{};
}
}  // namespace

size_t guess_noalloc_scratch_bytes() { return kNaBytes; }

Status GuessModel::predict_atoms_noalloc(const GuessAtoms& atoms, std::vector<float>& out) const {
    const int na = (int)atoms.z.size();
    const int ni = (int)atoms.inst_atom0.size() - 1;
    if (na > kNoAllocMaxAtoms || ni > kNoAllocMaxSystems)
        return Status::error(kErrInvalidInput, "the no-allocation prediction takes at most " +
                                                   std::to_string(kNoAllocMaxAtoms) + " atoms in " +
                                                   std::to_string(kNoAllocMaxSystems) + " systems");
    if (theta_.size() != (size_t)kGuessParams) return Status::error(kErrInvalidInput, "guess model has no parameters");
    if (!d_theta_ || !d_norm_)
        return Status::error(kErrInvalidInput, "the model's parameters are not on the device (no allocation allowed)");
    out.assign(2 * (size_t)na, 0.0f);
    if (na == 0) return Status{};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dirty_) {
        // new parameters of an already resident model: copies only
        GHIP(hipMemcpy(d_theta_, theta_.data(), kGuessParams * sizeof(float), hipMemcpyHostToDevice));
        GHIP(hipMemcpy(d_norm_, norm_.data(), 2 * kGuessFeat * sizeof(float), hipMemcpyHostToDevice));
        dirty_ = false;
    }
    unsigned char* base = nullptr;
    Status st = scratch_address(&base);
    if (!st.ok()) return st;
    int* z = (int*)(base + kNaZ);
    int* inst = (int*)(base + kNaInst);
    int* atom0 = (int*)(base + kNaAtom0);
    double* x = (double*)(base + kNaX);
    float* feat = (float*)(base + kNaFeat);
    float* y = (float*)(base + kNaOut);
    GHIP(hipMemcpy(z, atoms.z.data(), na * sizeof(int), hipMemcpyHostToDevice));
    GHIP(hipMemcpy(inst, atoms.inst.data(), na * sizeof(int), hipMemcpyHostToDevice));
    GHIP(hipMemcpy(atom0, atoms.inst_atom0.data(), (ni + 1) * sizeof(int), hipMemcpyHostToDevice));
    GHIP(hipMemcpy(x, atoms.x.data(), 3 * (size_t)na * sizeof(double), hipMemcpyHostToDevice));
    GHIP(guess_launch_features(0, na, z, x, inst, atom0, feat));
    GHIP(guess_launch_forward(0, na, d_theta_, d_norm_, feat, y));
    GHIP(hipMemcpy(out.data(), y, out.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < na; ++i)
        if (!atoms.usable[atoms.inst[i]]) out[2 * i] = out[2 * i + 1] = 0.0f;
    return Status{};
}

}  // namespace gfn1
