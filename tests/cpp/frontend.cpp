// Exercises the C++ drop-in header (include/gproshan/geodesics.h) the way the
// reference's callers use it (src/app_viewer.cpp:657-716).  Prints one line per
// check; argv[1] == "gpu" also runs PTP_GPU.
#include <gproshan/geodesics.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace gproshan;

int main(int argc, char **argv)
{
	const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
	// 20 x 20 planar grid, CCW triangles.
	const index_t n = 20;
	std::vector<vertex> V;
	std::vector<index_t> F;
	for (index_t r = 0; r < n; r++)
		for (index_t c = 0; c < n; c++) V.push_back({ c / (n - 1.0), r / (n - 1.0), 0.0 });
	for (index_t r = 0; r + 1 < n; r++)
		for (index_t c = 0; c + 1 < n; c++) {
			const index_t a = r * n + c, b = a + 1, d = a + n, e = d + 1;
			F.insert(F.end(), { a, b, e, a, e, d });
		}
	che mesh(V.data(), (index_t) V.size(), F.data(), (index_t) F.size() / 3);
	std::printf("mesh %zu vertices %zu faces %zu edges manifold %d\n", mesh.n_vertices(), mesh.n_faces(),
	            mesh.n_edges(), (int) mesh.is_manifold());
	const std::vector<index_t> src = { (n / 2) * n + n / 2 };

	geodesics cpu(&mesh, src, geodesics::PTP_CPU);
	double maxerr = 0;
	for (index_t v = 0; v < mesh.n_vertices(); v++) {
		const double eu = std::hypot(V[v].x - V[src[0]].x, V[v].y - V[src[0]].y);
		if (eu > 0.3) maxerr = std::fmax(maxerr, std::fabs(cpu[v] - eu) / eu);
	}
	std::printf("ptp_cpu far-field max rel err vs euclid %.4f sorted0 %u\n", maxerr, cpu(0));
	int ok = maxerr < 0.1 && cpu(0) == src[0];

	try {
		geodesics bad(&mesh, { 3, 3 }, geodesics::PTP_CPU);
		ok = 0;
	} catch (const ptp_error &e) {
		std::printf("duplicate sources -> status %d\n", e.status);
		ok &= e.status == GP_ERR_DUPLICATE_SOURCE;
	}

	if (gpu) {
		// PTP_GPU in its default FP64 and with GP_PREC_F32, written to
		// <prefix>.f64 / <prefix>.f32 (raw doubles) for the comparison with the
		// oracle in tests/test_frontend_cpp.py.
		std::vector<distance_t> ext(mesh.n_vertices()), ext32(mesh.n_vertices());
		geodesics g(&mesh, src, geodesics::PTP_GPU, ext.data());
		geodesics g32(&mesh, src, geodesics::PTP_GPU, ext32.data(), false, 0, INFINITY, GP_PREC_F32);
		if (argc > 2) {
			for (int q = 0; q < 2; q++) {
				const std::string path = std::string(argv[2]) + (q ? ".f32" : ".f64");
				FILE *f = std::fopen(path.c_str(), "wb");
				ok &= f != nullptr;
				if (f) {
					ok &= std::fwrite((q ? ext32 : ext).data(), sizeof(distance_t), ext.size(), f) == ext.size();
					std::fclose(f);
				}
			}
		}
		double d = 0;
		for (index_t v = 0; v < mesh.n_vertices(); v++) d = std::fmax(d, std::fabs(ext32[v] - ext[v]) / std::fmax(ext[v], 1e-30));
		std::printf("ptp_gpu f32 vs f64 max rel diff %.3e\n", d);
		ok &= d <= 1e-4;

		// Cluster labels of two sources at opposite corners: each corner's
		// own neighbourhood carries its label (1-based source index).
		{
			geodesics gc(&mesh, { 0, n * n - 1 }, geodesics::PTP_GPU, nullptr, true);
			std::printf("clusters %u %u %u %u\n", gc.clusters[0], gc.clusters[1], gc.clusters[n * n - 1],
			            gc.clusters[n * n - 2]);
			ok &= gc.clusters[0] == 1 && gc.clusters[1] == 1 && gc.clusters[n * n - 1] == 2 && gc.clusters[n * n - 2] == 2;
		}

		// Farthest-point sampling from the corner: the first sample added is
		// the opposite corner of the unit square.
		std::vector<index_t> samples = { 0 };
		double t = 0;
		const distance_t r = farthest_point_sampling_ptp_gpu(&mesh, samples, t, 8);
		std::printf("fps %zu samples, second %u, radio %.4f\n", samples.size(), samples[1], r);
		ok &= samples.size() == 8 && samples[1] == n * n - 1 && r > 0 && r < 1;
	}
	std::printf(ok ? "OK\n" : "FAILED\n");
	return ok ? 0 : 1;
}
