// dsx_gate_noalloc: the GPU seed gate of dsx_gate for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops). Every intermediate lives in one file-scope device array
// of this library (`scratch`, below); the caller passes only the outputs, in device memory, and a stream.
//
// Per call, enqueued on the caller's stream with no host synchronisation and no memset:
//   k_na_clear     zero the four gate counters (a kernel, so the call can be captured into a graph)
//   k_phase        y-phase filter over [start, start + count) into the phase-survivor list (gate.cuh:87-92)
//   k_na_clamp     the number of phase survivors the later stages take: all of them, or none if they
//                  overflow the scratch list (then counts[0] > DSX_NOALLOC_MAX_PHASE flags the call)
//   k_init         temperature octave-0 Perlin of both sides of each phase survivor (gate.cuh:93-97)
//   k_amp<0|2>     pre-gate; k_amp<1> full 512^2 gate (gate.cuh:205-346); grids sized for the scratch
//                  capacity, the kernels read the live count on the device and idle blocks exit at once
//   k_na_sort      rank sort of the full-gate survivors by seed index into the caller's arrays, and the
//                  counts (the order dsx_gate returns; its survivors are unique indices)
// The gate kernels are the ones dsx_gate runs, so survivors and amplitudes are bit-identical to it.
//
// Every call on a device shares that device's scratch array: calls that overlap on one device (two
// streams, two threads) race. The array keeps the last call's contents; everything a call reads is
// written by that call first.
#include "ctx.h"
#include <cstring>
#include <mutex>
#include <string>

namespace {

constexpr uint32_t MAXC = DSX_NOALLOC_MAX_COUNT;
constexpr uint32_t MAXP = DSX_NOALLOC_MAX_PHASE;

// Scratch layout (byte offsets): counters, phase-survivor indices, pre-gate survivors, full-gate survivor
// indices and amplitudes, then two noise sides per phase survivor.
constexpr size_t OFF_CNT = 0;                                        // u32[4]: phase, pre, full, taken
constexpr size_t OFF_K1 = 256;                                       // u32[MAXP]
constexpr size_t OFF_I2 = OFF_K1 + size_t(MAXP) * 4;                 // u32[MAXP]
constexpr size_t OFF_K3 = OFF_I2 + size_t(MAXP) * 4;                 // u32[MAXP]
constexpr size_t OFF_A3 = OFF_K3 + size_t(MAXP) * 4;                 // f32[MAXP]
constexpr size_t OFF_SIDES = (OFF_A3 + size_t(MAXP) * 4 + 255) & ~size_t(255);
constexpr size_t SCRATCH_BYTES = OFF_SIDES + 2 * size_t(MAXP) * sizeof(dsx::SideParams);
static_assert(SCRATCH_BYTES == DSX_NOALLOC_SCRATCH_BYTES, "scratch size differs from the documented one");

}  // namespace

// The scratch array. In host code `scratch` names a host-side stand-in, not device memory: its device
// address comes from hipGetSymbolAddress, per device (scratch_base below).
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

namespace {

__global__ void k_na_clear(uint32_t *cnt) {
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
}

__global__ void k_na_clamp(uint32_t *cnt) {
    if (threadIdx.x == 0) cnt[3] = cnt[0] <= MAXP ? cnt[0] : 0u;
}

// One thread per full-gate survivor: its rank among the survivors' (unique) seed indices is its output slot.
__global__ __launch_bounds__(256) void k_na_sort(const uint32_t *cnt, const uint32_t *k3, const float *a3, uint64_t start,
                                                 uint64_t *seeds, float *amps, uint64_t *kidx, uint64_t cap,
                                                 uint64_t *counts) {
    const uint32_t n = cnt[3] ? min(cnt[2], MAXP) : 0u;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0) {
        counts[0] = cnt[0];
        counts[1] = cnt[3] ? cnt[1] : 0u;
        counts[2] = n;
        counts[3] = uint64_t(n) < cap ? uint64_t(n) : cap;
    }
    if (i >= n) return;
    const uint32_t key = k3[i];
    uint32_t rank = 0;
    for (uint32_t j = 0; j < n; j++) rank += k3[j] < key ? 1u : 0u;
    if (rank >= cap) return;
    if (seeds) seeds[rank] = dsx::mix64(start + key);
    if (amps) amps[rank] = a3[i];
    if (kidx) kidx[rank] = key;
}

std::mutex g_mu;
constexpr int MAX_DEVICES = 64;
uint8_t *g_base[MAX_DEVICES] = {};   // device address of `scratch`, per device

// The device address of the scratch array on `device` (the current device), looked up once per device.
int scratch_base(int device, uint8_t **out) {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_base[device]) {
        hipDeviceProp_t prop;
        HIPCK(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            dsx::g_last_error = std::string("device is ") + prop.gcnArchName + ", library built for gfx950";
            return DSX_E_NODEVICE;
        }
        void *p = nullptr;
        HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            dsx::g_last_error = "hipGetSymbolAddress returned no address for the gate scratch array";
            return DSX_E_HIP;
        }
        g_base[device] = static_cast<uint8_t *>(p);
    }
    *out = g_base[device];
    return DSX_OK;
}

int gate_noalloc(int device, hipStream_t st, uint64_t start, uint64_t count, const dsx_gate_params *p, uint64_t *seeds,
                 float *amps, uint64_t *kidx, uint64_t cap, uint64_t *counts) {
    uint8_t *base = nullptr;
    const int rc = scratch_base(device, &base);
    if (rc != DSX_OK) return rc;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(base + OFF_CNT);
    uint32_t *k1 = reinterpret_cast<uint32_t *>(base + OFF_K1);
    uint32_t *i2 = reinterpret_cast<uint32_t *>(base + OFF_I2);
    uint32_t *k3 = reinterpret_cast<uint32_t *>(base + OFF_K3);
    float *a3 = reinterpret_cast<float *>(base + OFF_A3);
    dsx::SideParams *sides = reinterpret_cast<dsx::SideParams *>(base + OFF_SIDES);
    const dsx::GateKernel kpre = dsx::gate_kernel(4, p->box ? p->sig : 1.0f);
    const dsx::GateKernel kfull = dsx::gate_kernel(7, 2.0f);
    const double thr_pre = p->box ? double(p->pthr) : double(p->amp - p->margin);
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, st, cnt);
    HIPCK(hipGetLastError());
    if (count) {
        dsx::launch_phase(st, start, uint32_t(count), double(p->max_phase), k1, cnt, MAXP);
        HIPCK(hipGetLastError());
        hipLaunchKernelGGL(k_na_clamp, dim3(1), dim3(64), 0, st, cnt);
        HIPCK(hipGetLastError());
        dsx::launch_init(st, start, k1, nullptr, MAXP, sides, cnt + 3);
        HIPCK(hipGetLastError());
        dsx::launch_amp(st, p->box ? 0 : 2, sides, nullptr, MAXP, kpre, thr_pre, nullptr, i2, nullptr, cnt + 1, MAXP,
                        nullptr, cnt + 3);
        HIPCK(hipGetLastError());
        dsx::launch_amp(st, 1, sides, i2, MAXP, kfull, double(p->amp), k1, k3, a3, cnt + 2, MAXP, nullptr, cnt + 1);
        HIPCK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_na_sort, dim3(MAXP / 256), dim3(256), 0, st, cnt, k3, a3, start, seeds, amps, kidx, cap, counts);
    HIPCK(hipGetLastError());
    return DSX_OK;
}

}  // namespace

extern "C" int dsx_gate_noalloc(int device, void *stream, uint64_t start, uint64_t count, const dsx_gate_params *p,
                                uint64_t *seeds, float *amps, uint64_t *kidx, uint64_t cap, uint64_t *counts) {
    if (!dsx::valid_gate_params(p) || !counts) return DSX_E_ARG;
    if (count > MAXC) {
        dsx::g_last_error = "dsx_gate_noalloc: count exceeds DSX_NOALLOC_MAX_COUNT (2^22 seed indices per call)";
        return DSX_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= MAX_DEVICES) {
        dsx::g_last_error = "no HIP device";
        return DSX_E_NODEVICE;
    }
    int prev = -1;
    HIPCK(hipGetDevice(&prev));
    HIPCK(hipSetDevice(device));
    const int rc = gate_noalloc(device, static_cast<hipStream_t>(stream), start, count, p, seeds, amps, kidx, cap, counts);
    if (prev != device) (void)hipSetDevice(prev);
    return rc;
}
