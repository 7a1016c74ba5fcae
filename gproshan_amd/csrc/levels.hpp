// Multi-source BFS levels on the device and the vertices sorted by level (the
// rings of parallel_toplesets_propagation, che::compute_toplesets,
// src/che.cpp:546-593, as sets).  A frontier BFS over all workgroups (one
// launch per level, replayed from a captured graph, the host polling a done
// flag) lowers the levels from the seeded sources, stopping where the levels
// are already as low -- so seeding one more source into existing levels
// updates them incrementally -- and a counting sort by level rebuilds sorted[]
// and limits.  The order inside a ring is not the reference's discovery order
// (the window sweeps relax whole rings, so it changes no value).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gp {

struct LevelState;

struct Levels {
	uint32_t nv = 0;
	uint32_t *level = nullptr;    // [nv], NIL = unreached
	uint32_t *queue = nullptr;    // [nv]
	uint32_t *hist = nullptr;     // [nv + 1]
	uint32_t *sorted = nullptr;   // [nv]   vertices by level
	uint32_t *limits = nullptr;   // [nv + 2]
	uint32_t *nl = nullptr;       // |limits| (device)
	LevelState *ls = nullptr, *h_ls = nullptr;
	hipGraphExec_t graph = nullptr;
	hipStream_t stream = nullptr;   // borrowed
	const DeviceMesh *dm = nullptr;

	Levels() = default;
	Levels(const Levels &) = delete;
	Levels &operator=(const Levels &) = delete;
	~Levels();
	int create(const DeviceMesh *mesh, hipStream_t s);
	void clear();                                               // every level unreached
	void seed(const uint32_t *d_src, uint32_t first, uint32_t ns);   // d_src[first, ns) to level 0
	int expand();                                               // BFS to completion
	void sort();                                                // sorted, limits, nl on the device
};

}  // namespace gp
