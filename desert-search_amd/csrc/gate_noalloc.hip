// dsx_gate_noalloc: the GPU seed gate of dsx_gate for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops). Every intermediate lives in one file-scope device array
// of this library (`scratch`, below); the caller passes only the outputs, in device memory, and a stream.
//
// Per call, enqueued on the caller's stream with no host synchronisation and no memset: the gate chain of
// dsx_internal.h (enqueue_gate_chain, gate.hip) on the scratch array: counters cleared by a kernel, y-phase
// filter (gate.cuh:87-92), the count the later stages take (none if the phase survivors overflow the scratch
// list: counts[0] > DSX_NOALLOC_MAX_PHASE flags the call), Perlin init (gate.cuh:93-97), pre-gate and full
// gate (gate.cuh:205-346) with grids sized for the scratch capacity, and a rank sort of the survivors by seed
// index into the caller's arrays. These are the kernels dsx_gate runs, so results are bit-identical to it.
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
    dsx::GateChain C;
    C.cnt = reinterpret_cast<uint32_t *>(base + OFF_CNT);
    C.k1 = reinterpret_cast<uint32_t *>(base + OFF_K1);
    C.i2 = reinterpret_cast<uint32_t *>(base + OFF_I2);
    C.k3 = reinterpret_cast<uint32_t *>(base + OFF_K3);
    C.a3 = reinterpret_cast<float *>(base + OFF_A3);
    C.sides = reinterpret_cast<dsx::SideParams *>(base + OFF_SIDES);
    C.grid = MAXP;
    dsx::GateStages S;
    S.maxph = double(p->max_phase);
    S.thr_pre = p->box ? double(p->pthr) : double(p->amp - p->margin);
    S.amp = double(p->amp);
    S.mode_pre = p->box ? 0 : 2;
    S.full_fast = 0;
    S.kpre = dsx::gate_kernel(4, p->box ? p->sig : 1.0f);
    S.kfull = dsx::gate_kernel(7, 2.0f);
    HIPCK(dsx::enqueue_gate_chain(st, C, S, start, uint32_t(count), dsx::GateOut{seeds, amps, kidx, cap, counts}));
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
