// Mesh files into the CHE builder: OFF, OBJ and PLY readers with the
// reference's face conventions (src/che_off.cpp:28-79, src/che_obj.cpp:24-79,
// src/che_ply.cpp:25-156).
//
//   OFF   "OFF" / "COFF" (RGBA per vertex, skipped) / "NOFF" (normal per
//         vertex, skipped); counts; vertices; faces.  A file whose first face
//         is a quad is read as quads, each (a, b, c, d) split by index reuse
//         into (a, b, c) and (d, a, c) (che_off.cpp:62-72).
//   OBJ   "v x y z" and "f i j k [l]" lines (1-based, "i/t/n" forms read up to
//         the first '/'); a quad (a, b, c, d) becomes (a, b, d) and (b, c, d)
//         (che_obj.cpp:62-72); faces of other sizes are skipped as there.
//   PLY   ascii or binary_little_endian; x, y, z are the first three vertex
//         properties (che_ply.cpp:97-123), faces are "list <count> <index>"
//         triangles.  Binary faces read the count with its declared size (the
//         reference reads zero bytes there, che_ply.cpp:128-131).
// Differences: malformed input is an error (GP_ERR_INVALID with the line or
// field) instead of an assert or an out-of-bounds write; OFF files that mix
// face sizes and OFF/PLY faces with more than four corners are rejected.
#include "gp_internal.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace gp {

static bool ends_with(const std::string &s, const char *suf)
{
	const size_t n = std::strlen(suf);
	if (s.size() < n) return false;
	for (size_t q = 0; q < n; q++)
		if (std::tolower((unsigned char) s[s.size() - n + q]) != suf[q]) return false;
	return true;
}

static int read_off(std::istream &is, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	std::string head;
	uint64_t nv = 0, nf = 0, ne = 0;
	if (!(is >> head) || head.find("OFF") == std::string::npos)
		return fail(GP_ERR_INVALID, "off: missing OFF header");
	if (!(is >> nv >> nf >> ne)) return fail(GP_ERR_INVALID, "off: missing counts");
	if (nv >= SLOT_WRAP || nf > (1ull << 31)) return fail(GP_ERR_CAPACITY, "off: counts too large");
	const bool colour = head[0] == 'C', normal = head[0] == 'N';
	xyz.resize(3 * nv);
	for (uint64_t v = 0; v < nv; v++) {
		if (!(is >> xyz[3 * v] >> xyz[3 * v + 1] >> xyz[3 * v + 2]))
			return fail(GP_ERR_INVALID, "off: vertex %llu unreadable", (unsigned long long) v);
		double skip;
		for (int q = 0; q < (colour ? 4 : normal ? 3 : 0); q++)
			if (!(is >> skip)) return fail(GP_ERR_INVALID, "off: vertex %llu attributes unreadable", (unsigned long long) v);
	}
	faces.clear();
	faces.reserve(3 * nf);
	uint64_t first = 0;
	for (uint64_t f = 0; f < nf; f++) {
		uint64_t k = 0;
		if (!(is >> k)) return fail(GP_ERR_INVALID, "off: face %llu unreadable", (unsigned long long) f);
		if (f == 0) first = k;
		if (k != first || (k != 3 && k != 4))
			return fail(GP_ERR_INVALID, "off: face %llu has %llu corners (triangles or quads only, not mixed)",
			            (unsigned long long) f, (unsigned long long) k);
		uint32_t c[4];
		for (uint64_t q = 0; q < k; q++)
			if (!(is >> c[q])) return fail(GP_ERR_INVALID, "off: face %llu index unreadable", (unsigned long long) f);
		faces.insert(faces.end(), { c[0], c[1], c[2] });
		if (k == 4) faces.insert(faces.end(), { c[3], c[0], c[2] });
	}
	return GP_OK;
}

static int read_obj(std::istream &is, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	std::string line;
	uint64_t ln = 0;
	std::vector<int64_t> idx;
	while (std::getline(is, line)) {
		ln++;
		std::istringstream ss(line);
		std::string key;
		if (!(ss >> key)) continue;
		if (key == "v") {
			double x, y, z;
			if (!(ss >> x >> y >> z)) return fail(GP_ERR_INVALID, "obj: line %llu: bad vertex", (unsigned long long) ln);
			xyz.insert(xyz.end(), { x, y, z });
		} else if (key == "f") {
			idx.clear();
			std::string tok;
			while (ss >> tok) {
				const size_t slash = tok.find('/');
				const std::string head = tok.substr(0, slash);
				char *end = nullptr;
				const long long i = std::strtoll(head.c_str(), &end, 10);
				if (head.empty() || *end != '\0' || i <= 0)
					return fail(GP_ERR_INVALID, "obj: line %llu: bad face index '%s'", (unsigned long long) ln, tok.c_str());
				idx.push_back(i - 1);
			}
			if (idx.size() == 3) {
				faces.insert(faces.end(), { (uint32_t) idx[0], (uint32_t) idx[1], (uint32_t) idx[2] });
			} else if (idx.size() == 4) {
				faces.insert(faces.end(), { (uint32_t) idx[0], (uint32_t) idx[1], (uint32_t) idx[3] });
				faces.insert(faces.end(), { (uint32_t) idx[1], (uint32_t) idx[2], (uint32_t) idx[3] });
			}
		}
	}
	return GP_OK;
}

static size_t ply_size(const std::string &t)
{
	if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
	if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
	if (t == "int" || t == "uint" || t == "float" || t == "float32" || t == "int32" || t == "uint32") return 4;
	if (t == "double" || t == "float64") return 8;
	return 0;
}

static bool ply_float(const std::string &t) { return t == "float" || t == "float32" || t == "double" || t == "float64"; }

// Little-endian scalar of type t at p, as a double (indices and counts as
// unsigned where the type is).
static double ply_value(const unsigned char *p, const std::string &t)
{
	if (t == "float" || t == "float32") { float x; std::memcpy(&x, p, 4); return x; }
	if (t == "double" || t == "float64") { double x; std::memcpy(&x, p, 8); return x; }
	if (t == "char" || t == "int8") return (double) (int8_t) p[0];
	if (t == "uchar" || t == "uint8") return (double) p[0];
	if (t == "short" || t == "int16") { int16_t x; std::memcpy(&x, p, 2); return x; }
	if (t == "ushort" || t == "uint16") { uint16_t x; std::memcpy(&x, p, 2); return x; }
	if (t == "int" || t == "int32") { int32_t x; std::memcpy(&x, p, 4); return x; }
	uint32_t x;
	std::memcpy(&x, p, 4);
	return x;
}

static int read_ply(std::istream &is, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	std::string line, format, element;
	uint64_t nv = 0, nf = 0;
	std::vector<std::string> vtypes;          // vertex property types in order
	std::string cnt_t, idx_t;                 // face list types
	if (!std::getline(is, line) || line.rfind("ply", 0) != 0) return fail(GP_ERR_INVALID, "ply: missing header");
	bool end = false;
	while (std::getline(is, line)) {
		std::istringstream ss(line);
		std::string w;
		ss >> w;
		if (w == "end_header") { end = true; break; }
		if (w == "format") ss >> format;
		else if (w == "element") {
			ss >> element;
			if (element == "vertex") ss >> nv;
			else if (element == "face") ss >> nf;
		} else if (w == "property" && element == "vertex") {
			std::string t;
			ss >> t;
			if (!ply_size(t)) return fail(GP_ERR_INVALID, "ply: vertex property type '%s'", t.c_str());
			vtypes.push_back(t);
		} else if (w == "property" && element == "face") {
			std::string t;
			ss >> t;
			if (t == "list") ss >> cnt_t >> idx_t;
		}
	}
	if (!end) return fail(GP_ERR_INVALID, "ply: no end_header");
	if (vtypes.size() < 3 || !ply_float(vtypes[0])) return fail(GP_ERR_INVALID, "ply: x y z must be the first vertex properties");
	if (nv >= SLOT_WRAP || nf > (1ull << 31)) return fail(GP_ERR_CAPACITY, "ply: counts too large");
	if (nf && (!ply_size(cnt_t) || !ply_size(idx_t))) return fail(GP_ERR_INVALID, "ply: face list types");
	xyz.resize(3 * nv);
	faces.clear();
	faces.reserve(3 * nf);
	if (format == "ascii") {
		for (uint64_t v = 0; v < nv; v++) {
			if (!std::getline(is, line)) return fail(GP_ERR_INVALID, "ply: vertex %llu missing", (unsigned long long) v);
			std::istringstream ss(line);
			if (!(ss >> xyz[3 * v] >> xyz[3 * v + 1] >> xyz[3 * v + 2]))
				return fail(GP_ERR_INVALID, "ply: vertex %llu unreadable", (unsigned long long) v);
		}
		for (uint64_t f = 0; f < nf; f++) {
			if (!std::getline(is, line)) return fail(GP_ERR_INVALID, "ply: face %llu missing", (unsigned long long) f);
			std::istringstream ss(line);
			uint64_t k = 0;
			uint32_t c[3];
			if (!(ss >> k) || k != 3 || !(ss >> c[0] >> c[1] >> c[2]))
				return fail(GP_ERR_INVALID, "ply: face %llu is not a triangle", (unsigned long long) f);
			faces.insert(faces.end(), { c[0], c[1], c[2] });
		}
	} else if (format == "binary_little_endian") {
		size_t vbytes = 0;
		for (auto &t : vtypes) vbytes += ply_size(t);
		std::vector<unsigned char> buf(std::max<size_t>(vbytes, 8));
		const size_t s0 = ply_size(vtypes[0]), s1 = ply_size(vtypes[1]);
		for (uint64_t v = 0; v < nv; v++) {
			if (!is.read((char *) buf.data(), (std::streamsize) vbytes))
				return fail(GP_ERR_INVALID, "ply: vertex %llu truncated", (unsigned long long) v);
			xyz[3 * v] = ply_value(buf.data(), vtypes[0]);
			xyz[3 * v + 1] = ply_value(buf.data() + s0, vtypes[1]);
			xyz[3 * v + 2] = ply_value(buf.data() + s0 + s1, vtypes[2]);
		}
		const size_t cs = ply_size(cnt_t), is_ = ply_size(idx_t);
		for (uint64_t f = 0; f < nf; f++) {
			if (!is.read((char *) buf.data(), (std::streamsize) cs))
				return fail(GP_ERR_INVALID, "ply: face %llu truncated", (unsigned long long) f);
			const double k = ply_value(buf.data(), cnt_t);
			if (k != 3) return fail(GP_ERR_INVALID, "ply: face %llu is not a triangle", (unsigned long long) f);
			for (int q = 0; q < 3; q++) {
				if (!is.read((char *) buf.data(), (std::streamsize) is_))
					return fail(GP_ERR_INVALID, "ply: face %llu truncated", (unsigned long long) f);
				const double i = ply_value(buf.data(), idx_t);
				if (i < 0) return fail(GP_ERR_INVALID, "ply: face %llu negative index", (unsigned long long) f);
				faces.push_back((uint32_t) i);
			}
		}
	} else {
		return fail(GP_ERR_INVALID, "ply: format '%s' not supported", format.c_str());
	}
	return GP_OK;
}

int read_mesh(const char *path, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	if (!path) return fail(GP_ERR_INVALID, "mesh file: null path");
	const std::string p(path);
	std::ifstream is(p, std::ios::binary);
	if (!is.good()) return fail(GP_ERR_INVALID, "mesh file: cannot open '%s'", path);
	xyz.clear();
	faces.clear();
	if (ends_with(p, ".off")) return read_off(is, xyz, faces);
	if (ends_with(p, ".obj")) return read_obj(is, xyz, faces);
	if (ends_with(p, ".ply")) return read_ply(is, xyz, faces);
	return fail(GP_ERR_INVALID, "mesh file: unknown extension of '%s' (.off, .obj, .ply)", path);
}

}  // namespace gp
