// Multi-source BFS levels on the device and the vertices sorted by level (the
// rings of parallel_toplesets_propagation, che::compute_toplesets,
// src/che.cpp:546-593, as sets).  A frontier BFS over all workgroups (one
// launch per level, replayed from a graph, the host polling a done
// flag) lowers the levels from the seeded sources, stopping where the levels
// are already as low -- so seeding one more source into existing levels
// updates them incrementally -- and a counting sort by level and, inside a
// level, by slot count (longest fans first) rebuilds sorted[], limits and the
// slot rows of each 64-position chunk.  The order inside a ring is not the
// reference's discovery order (the window sweeps relax whole rings, so it
// changes no value).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gp {

struct LevelState;

constexpr uint32_t LEVEL_KEYS = 8;   // slot classes per level in the sort key

struct Levels {
	uint32_t nv = 0;
	uint32_t *level = nullptr;    // [nv], NIL = unreached
	uint32_t *queue = nullptr;    // [nv]
	uint32_t *hist = nullptr;     // [nv * LEVEL_KEYS + 1]  key counts, then cursors
	uint32_t *sorted = nullptr;   // [nv]   vertices by level, longest fans first in a level
	uint32_t *limits = nullptr;   // [nv + 2]
	uint32_t *nl = nullptr;       // |limits| (device)
	uint32_t *cslots = nullptr;   // [nchunks() + 1] slot rows of each 64-position chunk of sorted
	uint32_t *coff = nullptr;     // [nchunks() + 1] their exclusive prefix (chunk c at row coff[c])
	LevelState *ls = nullptr, *h_ls = nullptr;   // h_ls: two pinned snapshots
	hipEvent_t ev[2] = { nullptr, nullptr };
	hipGraphExec_t graph = nullptr;   // null when graphs are disabled
	uint32_t lv_grid = 1;
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
	void sort();                                                // sorted, limits, nl, coff on the device
	uint32_t nchunks() const { return (nv + 63) / 64; }
};

}  // namespace gp
