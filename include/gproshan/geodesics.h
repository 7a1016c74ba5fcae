/*
 * gproshan/geodesics.h -- C++ drop-in for gproshan's PTP front end, header
 * only, over the C ABI of libgproshan_ptp.so (include/gproshan_ptp.h).
 *
 * Mirrors, with the reference's names, argument meaning and ownership rules:
 *   class che (construction and the accessors PTP uses)   include/che.h:27-130
 *   che::compute_toplesets                                include/che.h:107
 *   struct ptp_out_t, struct toplesets_t                  include/geodesics_ptp.h:18-30
 *   parallel_toplesets_propagation_{gpu,coalescence_gpu,cpu,coalescence_cpu}
 *                                                         include/geodesics_ptp.h:34-40
 *   normalize_ptp                                         include/geodesics_ptp.h:46
 *   farthest_point_sampling_ptp_gpu                       include/geodesics_ptp.h:42
 *   che_off / che_obj / che_ply (file readers)            src/che_off.cpp, che_obj.cpp, che_ply.cpp
 *   class geodesics (PTP_CPU / PTP_GPU options)           include/geodesics.h:18-72
 * Differences: errors throw gproshan::ptp_error (the reference asserts or
 * reads out of bounds); FM and the heat method are other algorithms and
 * throw when requested; cluster labels come from PTP_GPU only (the
 * reference's CPU rule is a racy in-place update) and are NIL where
 * unreached.
 */
#ifndef GPROSHAN_GEODESICS_H
#define GPROSHAN_GEODESICS_H

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gproshan_ptp.h"

namespace gproshan {

typedef uint32_t index_t;        // include/include.h:19
typedef double real_t;           // include/include.h:21-25 (default build)
typedef real_t distance_t;
constexpr index_t NIL = GP_NIL;  // include/include.h:12

struct vertex {
	real_t x, y, z;
};

class ptp_error : public std::runtime_error {
  public:
	int status;
	ptp_error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int rc)
{
	if (rc != GP_OK) throw ptp_error(rc, gp_last_error());
}

class che {
  public:
	static const size_t P = 3;

	che(const vertex *vertices, const index_t &n_v, const index_t *faces, const index_t &n_f)
	{
		check(gp_che_create(&vertices[0].x, n_v, faces, n_f, &h_));
		attach();
	}
	// che_off / che_obj / che_ply(file): the format follows the extension.
	explicit che(const std::string &file)
	{
		check(gp_che_read(file.c_str(), &h_));
		attach();
	}
	~che() { gp_che_destroy(h_); }
	che(const che &) = delete;
	che &operator=(const che &) = delete;

	const size_t &n_vertices() const { return nv_; }
	const size_t &n_faces() const { return nf_; }
	const size_t &n_half_edges() const { return nhe_; }
	const size_t &n_edges() const { return ne_; }
	bool is_manifold() const { return manifold_; }
	const index_t &vt(const index_t &he) const { return vt_.at(he); }
	const index_t &ot(const index_t &he) const { return ot_.at(he); }
	const index_t &evt(const index_t &v) const { return evt_.at(v); }

	void link(std::vector<index_t> &l, const index_t &v) const
	{
		uint32_t n = 0;
		std::vector<index_t> buf(256);
		check(gp_che_link(h_, v, buf.data(), (uint32_t) buf.size(), &n));
		l.insert(l.end(), buf.begin(), buf.begin() + n);
	}

	// Same contract as the reference: the caller provides toplesets[n] and
	// sorted[n]; limits is appended to.
	void compute_toplesets(index_t *&toplesets, index_t *&sorted, std::vector<index_t> &limits,
	                       const std::vector<index_t> &sources, const index_t &k = NIL) const
	{
		if (sources.empty()) return;
		std::vector<uint32_t> lim(nv_ + 2);
		uint32_t nl = 0;
		check(gp_compute_toplesets_k(h_, sources.data(), (uint32_t) sources.size(), k, toplesets, sorted, lim.data(),
		                             (uint32_t) lim.size(), &nl));
		limits.insert(limits.end(), lim.begin(), lim.begin() + nl);
	}

	gp_che *handle() const { return h_; }

  private:
	void attach()
	{
		int man = 0;
		uint32_t nv, nf, nhe, ne;
		check(gp_che_sizes(h_, &nv, &nf, &nhe, &ne, &man));
		nv_ = nv; nf_ = nf; nhe_ = nhe; ne_ = ne; manifold_ = man != 0;
		vt_.resize(nhe_); ot_.resize(nhe_); evt_.resize(nv_);
		check(gp_che_tables(h_, vt_.data(), ot_.data(), evt_.data()));
	}

	gp_che *h_ = nullptr;
	size_t nv_ = 0, nf_ = 0, nhe_ = 0, ne_ = 0;
	bool manifold_ = true;
	std::vector<index_t> vt_, ot_, evt_;
};

// The reference's reader classes (include/che_off.h, che_obj.h, che_ply.h).
struct che_off : che {
	explicit che_off(const std::string &file) : che(file) {}
};
struct che_obj : che {
	explicit che_obj(const std::string &file) : che(file) {}
};
struct che_ply : che {
	explicit che_ply(const std::string &file) : che(file) {}
};

struct ptp_out_t {
	distance_t *dist;
	index_t *clusters;
	ptp_out_t(distance_t *const &d, index_t *const &c = nullptr) : dist(d), clusters(c) {}
};

struct toplesets_t {
	const std::vector<index_t> &limits;
	const index_t *const &index;
};

inline double parallel_toplesets_propagation_gpu(const ptp_out_t &out, che *mesh, const std::vector<index_t> &sources,
                                                 const toplesets_t &t)
{
	double s = 0;
	check(gp_parallel_toplesets_propagation_gpu(mesh->handle(), sources.data(), (uint32_t) sources.size(),
	                                            t.limits.data(), (uint32_t) t.limits.size(), t.index, out.dist,
	                                            out.clusters, &s));
	return s;
}

inline double parallel_toplesets_propagation_coalescence_gpu(const ptp_out_t &out, che *mesh,
                                                             const std::vector<index_t> &sources,
                                                             const toplesets_t &t, const bool & = 1)
{
	double s = 0;
	check(gp_parallel_toplesets_propagation_coalescence_gpu(mesh->handle(), sources.data(), (uint32_t) sources.size(),
	                                                        t.limits.data(), (uint32_t) t.limits.size(), t.index,
	                                                        out.dist, out.clusters, &s));
	return s;
}

inline void parallel_toplesets_propagation_cpu(const ptp_out_t &out, che *mesh, const std::vector<index_t> &sources,
                                               const toplesets_t &t)
{
	if (out.clusters) throw ptp_error(GP_ERR_INVALID, "clusters are not produced");
	check(gp_parallel_toplesets_propagation_cpu(mesh->handle(), sources.data(), (uint32_t) sources.size(),
	                                            t.limits.data(), (uint32_t) t.limits.size(), t.index, out.dist));
}

inline void parallel_toplesets_propagation_coalescence_cpu(const ptp_out_t &out, che *mesh,
                                                           const std::vector<index_t> &sources, const toplesets_t &t)
{
	if (out.clusters) throw ptp_error(GP_ERR_INVALID, "clusters are not produced");
	check(gp_parallel_toplesets_propagation_coalescence_cpu(mesh->handle(), sources.data(), (uint32_t) sources.size(),
	                                                        t.limits.data(), (uint32_t) t.limits.size(), t.index,
	                                                        out.dist));
}

inline void normalize_ptp(distance_t *dist, const size_t &n) { check(gp_normalize_ptp(dist, (uint32_t) n)); }

// farthest_point_sampling_ptp_gpu (include/geodesics_ptp.h:42): appends to
// samples, returns the last read maximum distance, time_fps in seconds.
inline distance_t farthest_point_sampling_ptp_gpu(che *mesh, std::vector<index_t> &samples, double &time_fps,
                                                  size_t n, distance_t radio = 0)
{
	const size_t nv = mesh->n_vertices();
	const size_t want = n >= nv ? nv / 2 : n;
	const size_t n0 = samples.size();
	samples.resize(std::max(n0, want));
	uint32_t cnt = 0;
	double r = 0;
	check(gp_farthest_point_sampling_ptp_gpu(mesh->handle(), samples.data(), (uint32_t) n0, (uint32_t) samples.size(),
	                                         (uint64_t) n, radio, GP_PREC_F64, &cnt, &r, &time_fps));
	samples.resize(cnt);
	return r;
}

class geodesics {
  public:
	// Numbering of the reference's CUDA build (include/geodesics.h:21-28).
	enum option_t { FM, PTP_GPU, HEAT_FLOW_GPU, PTP_CPU, HEAT_FLOW };

	index_t *clusters = nullptr;

	// precision (no reference counterpart): the GPU arithmetic of PTP_GPU.
	// The default, GP_PREC_F64, is the reference's default build
	// (include/config.h:7-8) and returns its bits; GP_PREC_F32 is the FP32
	// bulk update with the FP64 fallback (within 1e-4 of it).
	geodesics(che *mesh, const std::vector<index_t> &sources, const option_t &opt = FM,
	          distance_t *const &e_dist = nullptr, const bool &cluster = false, const size_t &n_iter = 0,
	          const distance_t &radio = INFINITY, const int &precision = GP_PREC_F64)
		: n_vertices_(mesh->n_vertices())
	{
		(void) n_iter; (void) radio;                      // ignored by PTP in the reference too
		if (n_vertices_ == 0) throw ptp_error(GP_ERR_INVALID, "mesh has no vertices");
		if (sources.empty()) throw ptp_error(GP_ERR_INVALID, "no sources");
		if (cluster && opt != PTP_GPU) throw ptp_error(GP_ERR_INVALID, "cluster labels are produced by PTP_GPU only");
		if (opt != PTP_CPU && opt != PTP_GPU) throw ptp_error(GP_ERR_INVALID, "only PTP_CPU and PTP_GPU are provided");
		free_dist_ = e_dist == nullptr;
		dist_ = free_dist_ ? new distance_t[n_vertices_] : e_dist;
		sorted_index_ = new index_t[n_vertices_];
		std::memset(sorted_index_, 0xFF, n_vertices_ * sizeof(index_t));
		for (size_t v = 0; v < n_vertices_; v++) dist_[v] = INFINITY;
		try {
			if (opt == PTP_CPU) {
				std::vector<index_t> top(n_vertices_), limits;
				index_t *tp = top.data();
				mesh->compute_toplesets(tp, sorted_index_, limits, sources);
				parallel_toplesets_propagation_coalescence_cpu({ dist_ }, mesh, sources, { limits, sorted_index_ });
			} else if (cluster) {
				// The GPU drivers with their cluster output (src/geodesics.cpp:225-240).
				clusters = new index_t[n_vertices_];
				std::vector<index_t> top(n_vertices_), limits;
				index_t *tp = top.data();
				mesh->compute_toplesets(tp, sorted_index_, limits, sources);
				if (sources.size() == 1)
					parallel_toplesets_propagation_coalescence_gpu({ dist_, clusters }, mesh, sources,
					                                               { limits, sorted_index_ });
				else
					parallel_toplesets_propagation_gpu({ dist_, clusters }, mesh, sources, { limits, sorted_index_ });
			} else {
				// The reference runs the coalescence driver for one source and
				// the sorted[] driver for several (src/geodesics.cpp:232-235);
				// both give the same field.  Here both run ring-numbered sweeps
				// whose rings and tables are built on the device (host build
				// for fans past the register-held slots).
				gp_ptp_opts o;
				gp_ptp_opts_default(&o);
				o.precision = precision;
				o.coalesce = 1;
				o.device_setup = 1;
				check(gp_ptp(mesh->handle(), sources.data(), (uint32_t) sources.size(), &o, dist_, sorted_index_,
				             nullptr));
			}
		} catch (...) {
			release();
			throw;
		}
	}

	virtual ~geodesics() { release(); }
	geodesics(const geodesics &) = delete;
	geodesics &operator=(const geodesics &) = delete;

	const distance_t &operator[](const index_t &i) const { return dist_[at(i)]; }
	const index_t &operator()(const index_t &i) const { return sorted_index_[at(i)]; }
	const size_t &n_sorted_index() const { return n_sorted_; }
	const index_t &farthest() const
	{
		if (!n_sorted_) throw ptp_error(GP_ERR_INVALID, "farthest() needs a sorted run; PTP leaves n_sorted = 0");
		return sorted_index_[n_sorted_ - 1];
	}
	const distance_t &radio() const { return dist_[farthest()]; }
	void copy_sorted_index(index_t *indexes, const size_t &n) const
	{
		if (n > n_sorted_) throw ptp_error(GP_ERR_INVALID, "n exceeds n_sorted_index()");
		std::memcpy(indexes, sorted_index_, n * sizeof(index_t));
	}
	void normalize() { normalize_ptp(dist_, n_vertices_); }   // n_sorted == 0: src/geodesics.cpp:77-83

  private:
	size_t at(const index_t &i) const
	{
		if (i >= n_vertices_) throw std::out_of_range("vertex index out of range");
		return i;
	}
	void release()
	{
		if (free_dist_) delete[] dist_;
		delete[] sorted_index_;
		delete[] clusters;
		dist_ = nullptr;
		sorted_index_ = nullptr;
		clusters = nullptr;
	}

	distance_t *dist_ = nullptr;
	index_t *sorted_index_ = nullptr;
	size_t n_vertices_ = 0, n_sorted_ = 0;
	bool free_dist_ = true;
};

}  // namespace gproshan

#endif  // GPROSHAN_GEODESICS_H
