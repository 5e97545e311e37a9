// dgx_noalloc.hip -- dgx_ssp_rk3_step_noalloc: the explicit SSP-RK3 step of dgx_ssp_rk3_step with every buffer
// its kernels use beyond the caller's state (the stage-0 copy, the residual, the max-norm slot and the limiter
// flag) in one file-scope device array instead of memory the library allocates.  For callers that cannot call
// hipMalloc (graph capture, tight time loops, allocation-free harnesses).
//
// The array is per device (each GPU that runs this library holds its own copy, at its own address, until the
// process ends).  Host code never passes the array's name to a kernel: it is a host-side stand-in.  The device
// address is looked up once per device with hipGetSymbolAddress (from dgx_create, with the context's device
// current) and used everywhere host code needs the memory.  Every call on one GPU shares that GPU's array and
// leaves its contents behind; two calls that overlap on one GPU (different streams or host threads) race.
#include <mutex>

#include "dgx_internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per device
constexpr size_t kHead = 256;                         // max-norm slot at 0, limiter flag at 8
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_ptr[kMaxDevices] = {};
std::mutex g_mu;

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

namespace dgx {

hipError_t noalloc_scratch(int dev, void** ptr)
{
	if (dev < 0 || dev >= kMaxDevices || !ptr) return hipErrorInvalidValue;
	std::lock_guard<std::mutex> lock(g_mu);
	if (!g_ptr[dev]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return e;
		g_ptr[dev] = p;
	}
	*ptr = g_ptr[dev];
	return hipSuccess;
}

}  // namespace dgx

extern "C" dgx_status dgx_ssp_rk3_step_noalloc(dgx_ctx* h, double* u, double dt, int positivity, double* max_abs_r)
{
	if (!h || !u) return DGX_ERR_ARG;
	dgx::Ctx& c = h->c;
	if (c.hp) return DGX_ERR_UNSUPPORTED;                            // variable order: per-group buffers
	if (!c.mesh_ready) return DGX_ERR_NO_MESH;
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return DGX_ERR_UNSUPPORTED;   // its gradient traces are not in scratch
	const size_t n = (size_t)c.E * c.Np * c.nv;
	if (kHead + 2 * round256(sizeof(double) * n) > kScratchBytes) return DGX_ERR_ARG;   // beyond the 64 MiB array
	if (!c.noalloc_scratch) return DGX_ERR_HIP;                     // the look-up at dgx_create failed
	unsigned char* base = static_cast<unsigned char*>(c.noalloc_scratch);
	auto* dmax = reinterpret_cast<unsigned long long*>(base);
	int* flag = reinterpret_cast<int*>(base + 8);
	double* u0 = reinterpret_cast<double*>(base + kHead);
	double* r = reinterpret_cast<double*>(base + kHead + round256(sizeof(double) * n));
	const bool pos = positivity && c.cfg.pde != DGX_PDE_ADVECTION;
	// the array keeps the previous call's contents: clear the flag and the max-norm slot with fill kernels
	if (pos && dgx::launch_fill_i32(c.stream, flag, 1, 0) != hipSuccess) return DGX_ERR_HIP;
	for (int rk = 0; rk < 3; ++rk) {
		if (dgx::launch_fill_u64(c.stream, dmax, 1, 0ull) != hipSuccess ||
		    dgx::launch_rhs64(c, u, r, dgx::MODE_RHS_MINV, dmax) != hipSuccess ||
		    dgx::launch_rk_stage_buf(c, u, u0, r, dt, rk) != hipSuccess)
			return DGX_ERR_HIP;
		if (pos && dgx::launch_positivity(c, u, flag) != hipSuccess) return DGX_ERR_HIP;
	}
	if (!max_abs_r && !pos) return DGX_OK;
	unsigned long long bits = 0;
	int bad = 0;
	if ((max_abs_r && hipMemcpyAsync(&bits, dmax, sizeof(bits), hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
	    (pos && hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
	    hipStreamSynchronize(c.stream) != hipSuccess)
		return DGX_ERR_HIP;
	if (bad) return DGX_ERR_POSITIVITY;
	if (max_abs_r) {
		double v;
		static_assert(sizeof(v) == sizeof(bits), "max-norm slot");
		__builtin_memcpy(&v, &bits, sizeof(v));
		*max_abs_r = v;
		if (v != v || v > 1.7e308) return DGX_ERR_NONFINITE;
	}
	return DGX_OK;
}
