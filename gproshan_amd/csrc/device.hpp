// Device-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "gp_internal.hpp"

namespace gp {

#define GP_HIP(call)                                                                  \
	do {                                                                              \
		hipError_t e_ = (call);                                                       \
		if (e_ != hipSuccess)                                                         \
			return fail(GP_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
	} while (0)

static_assert(sizeof(Geo32) == sizeof(float4), "Geo32 must match float4");

struct MasterLayout;  // ptp_gpu.hip

struct DeviceMesh {
	uint32_t nv = 0, nfan = 0;
	uint32_t *fan_off = nullptr;
	uint32_t *fan = nullptr;   // fan lists (vertex ids)
	double *pos = nullptr;     // 3*nv
	uint8_t *closed = nullptr; // 1 when the fan of v is closed (link = first m entries)
	float h = 1.f, inv_h = 1.f;
	MasterLayout *master[2] = { nullptr, nullptr };   // vertex-major slot tables (FP32, FP64), built on demand
};

// Uploads the fan tables and FP64 positions of m once (m.dev caches them).
int upload_mesh(Che &m);


}  // namespace gp
