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

struct DeviceMesh {
	uint32_t nv = 0, nfan = 0, n_flagged = 0;
	uint32_t *fan_off = nullptr;
	uint32_t *fan = nullptr;   // ids; bit 31 marks triangle (k, k+1) for the FP64 update
	float4 *geo = nullptr;     // per entry: q0, g, det, S in units of h
	double *pos = nullptr;     // 3*nv
	uint8_t *closed = nullptr; // 1 when the fan of v is closed (link = first m entries)
	float h = 1.f, inv_h = 1.f;
};

// Uploads the fan tables and FP32 geometry of m once (m.dev caches them).
int upload_mesh(Che &m);


}  // namespace gp
