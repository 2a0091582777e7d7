// cfg_nonlinear_noalloc: one NSE::nonlinear evaluation on device-resident
// fields that allocates nothing (include/cfgpu.h). It has no counterpart in
// the reference: it is for callers that cannot allocate device memory (code
// under HIP graph capture set-up, tight serving loops, harnesses that forbid
// hipMalloc).
//
// Everything the nonlinear kernels need beyond the caller's u and f -- the
// member table, the FFT twiddles, the y operators, the base-flow values and
// the work buffers of the transforms -- lives in one file-scope device array,
// `scratch`, of kScratchBytes per device. Its device address is looked up
// once per device with hipGetSymbolAddress and kept; host code only ever uses
// that address. Every call on a device shares the array and leaves its
// contents behind: calls that overlap on one device race.
#include "cfg_step.hpp"

using namespace cfg;

namespace {

constexpr size_t kScratchBytes = size_t(48) << 20;  // 48 MiB per device
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

double* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

double* scratch_ptr() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(CFG_ERR_DEVICE, "device ordinal out of range");
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(CFG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

// Host tables of one (grid, dealiasing, base flow): twiddles, y operators and
// the base-flow values, in the order they are copied into the scratch array.
struct HostNL {
    Geom g{};
    FFTPlan px{}, pz{};
    std::vector<double> twx, twz, T, TD, Ti, DTi, base;
};

// A region of `n` doubles at a 256-byte boundary
size_t carve(size_t& off, size_t n) {
    const size_t at = off;
    off = (off + n + 31) & ~size_t(31);
    return at;
}

}  // namespace

extern "C" int cfg_nonlinear_noalloc(const cfg_grid* grid, const cfg_flags* flags, const double* u_dev,
                                     double* f_dev, void* stream) {
    return guarded([&] {
        if (!grid || !flags || !u_dev || !f_dev) throw Error(CFG_ERR_ARG, "null argument");
        check_flags(*flags);
        check_grid(*grid);
        const bool dealias = flags->dealiasing == CFG_DealiasXZ;
        HostNL h;
        h.g = make_geom(*grid, dealias);
        const Geom& g = h.g;
        // scratch layout, in doubles
        size_t off = 0;
        const size_t o_tab = carve(off, (sizeof(NLMember) + 7) / 8);
        const size_t o_twx = carve(off, 2 * (size_t)g.Nx);
        const size_t o_twz = carve(off, 2 * (size_t)g.Nz);
        const size_t nyy = (size_t)g.Ny * g.Ny;
        const size_t o_T = carve(off, nyy), o_TD = carve(off, nyy), o_Ti = carve(off, nyy), o_DTi = carve(off, nyy);
        const size_t o_base = carve(off, 4 * (size_t)g.Ny);
        const size_t ny = (size_t)nlwork_y_doubles(g), nx = (size_t)nlwork_x_doubles(g);
        const size_t o_yp = carve(off, ny), o_fo = carve(off, ny), o_xp = carve(off, nx), o_zo = carve(off, nx);
        if (off > kScratchDoubles)
            throw Error(CFG_ERR_ARG, "grid " + std::to_string(g.Nx) + "x" + std::to_string(g.Ny) + "x" +
                                         std::to_string(g.Nz) + " needs " + std::to_string(off * 8) +
                                         " scratch bytes; cfg_nonlinear_noalloc holds " +
                                         std::to_string(kScratchBytes));
        init_device_limits();
        double* S = scratch_ptr();
        hipStream_t st = static_cast<hipStream_t>(stream);

        // host tables (the same expressions as the allocating path's Geometry and FlowModel)
        h.px = fft_plan(g.Nx, h.twx);
        h.pz = fft_plan(g.Nz, h.twz);
        YTables yt = y_tables(g.Ny, g.a, g.b);
        std::vector<double> U, W;
        base_from_flags(g, *flags, U, W);
        h.base = base_values_host(g, yt.tcheb, U, W);
        h.px.tw = reinterpret_cast<const double2*>(S + o_twx);
        h.pz.tw = reinterpret_cast<const double2*>(S + o_twz);
        const NLMember m{u_dev, f_dev, S + o_base, flags->Vsuck, flags->rotation};
        auto put = [&](size_t at, const void* src, size_t bytes) {
            HIPCHK(hipMemcpyAsync(S + at, src, bytes, hipMemcpyHostToDevice, st));
        };
        put(o_tab, &m, sizeof m);
        put(o_twx, h.twx.data(), h.twx.size() * 8);
        put(o_twz, h.twz.data(), h.twz.size() * 8);
        put(o_T, yt.T.data(), nyy * 8);
        put(o_TD, yt.TD.data(), nyy * 8);
        put(o_Ti, yt.Ti.data(), nyy * 8);
        put(o_DTi, yt.DTi.data(), nyy * 8);
        put(o_base, h.base.data(), h.base.size() * 8);
        int method;  // NSE::nonlinear's form for one call (nse.cpp:12-91)
        switch (flags->nonlinearity) {
            case CFG_Rotational: method = NL_ROT; break;
            case CFG_Convection: method = NL_CONV; break;
            case CFG_Divergence: method = NL_DIV; break;
            case CFG_SkewSymmetric: method = NL_SKEW; break;
            case CFG_Alternating: method = NL_DIV; break;
            default: method = NL_CONV; break;
        }
        const YMats ym{S + o_T, S + o_TD, S + o_Ti, S + o_DTi};
        const NLWork w{S + o_yp, S + o_xp, S + o_zo, S + o_fo};
        launch_nonlinear(g, ym, h.px, h.pz, method, 1, reinterpret_cast<const NLMember*>(S + o_tab),
                         flags->rotation != 0.0, w, NLTune{}, st, nullptr);
        HIPCHK(hipGetLastError());
        // the host tables above are read by the copies: wait for them (and the result) before returning
        HIPCHK(hipStreamSynchronize(st));
    });
}
