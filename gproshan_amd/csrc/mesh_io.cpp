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

#include <charconv>
#include <clocale>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <locale.h>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

namespace gp {

// Mesh files are read whole and parsed from memory: numbers with
// std::from_chars, which rounds decimal text to the nearest double as the
// stream extraction of the first version of these readers did (a leading '+'
// is accepted as there; text that is not a finite number is an error, as the
// stream's failure was).
struct Text {
	const char *p, *e;

	static bool space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
	void skip()
	{
		while (p < e && space(*p)) p++;
	}
	bool word(std::string &w)
	{
		skip();
		const char *b = p;
		while (p < e && !space(*p)) p++;
		w.assign(b, p);
		return p > b;
	}
	template <typename T>
	bool num(T &x)
	{
		skip();
		const char *b = p;
		if (e - b > 1 && *b == '+' && b[1] != '+' && b[1] != '-') b++;
		const auto r = std::from_chars(b, e, x);
		if constexpr (std::is_floating_point_v<T>) {
			if (r.ec == std::errc::result_out_of_range) {
				// Underflow reads as the nearest value (a signed 0), as the
				// stream took it; overflow stays an error.  strtod_l in the
				// "C" locale, whatever locale the host program has set.
				static const locale_t c_locale = newlocale(LC_ALL_MASK, "C", (locale_t) 0);
				if (!c_locale) return false;
				const double v = strtod_l(std::string(b, r.ptr).c_str(), nullptr, c_locale);
				if (!std::isfinite(v)) return false;
				x = (T) v;
				p = r.ptr;
				return true;
			}
		}
		if (r.ec != std::errc()) return false;
		if constexpr (std::is_floating_point_v<T>)
			if (!std::isfinite(x)) return false;
		p = r.ptr;
		return true;
	}
	// The rest of the current line (without its '\n'), and the position moved
	// past it; false at the end of the text.
	bool line(Text &l)
	{
		if (p >= e) return false;
		const char *n = static_cast<const char *>(std::memchr(p, '\n', (size_t) (e - p)));
		l = { p, n ? n : e };
		p = n ? n + 1 : e;
		return true;
	}
};

static bool ends_with(const std::string &s, const char *suf)
{
	const size_t n = std::strlen(suf);
	if (s.size() < n) return false;
	for (size_t q = 0; q < n; q++)
		if (std::tolower((unsigned char) s[s.size() - n + q]) != suf[q]) return false;
	return true;
}

static int read_off(Text &is, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	std::string head;
	uint64_t nv = 0, nf = 0, ne = 0;
	if (!is.word(head) || head.find("OFF") == std::string::npos)
		return fail(GP_ERR_INVALID, "off: missing OFF header");
	if (!(is.num(nv) && is.num(nf) && is.num(ne))) return fail(GP_ERR_INVALID, "off: missing counts");
	if (nv >= SLOT_WRAP || nf > (1ull << 31)) return fail(GP_ERR_CAPACITY, "off: counts too large");
	const bool colour = head[0] == 'C', normal = head[0] == 'N';
	xyz.resize(3 * nv);
	for (uint64_t v = 0; v < nv; v++) {
		if (!(is.num(xyz[3 * v]) && is.num(xyz[3 * v + 1]) && is.num(xyz[3 * v + 2])))
			return fail(GP_ERR_INVALID, "off: vertex %llu unreadable", (unsigned long long) v);
		double skip;
		for (int q = 0; q < (colour ? 4 : normal ? 3 : 0); q++)
			if (!is.num(skip)) return fail(GP_ERR_INVALID, "off: vertex %llu attributes unreadable", (unsigned long long) v);
	}
	faces.clear();
	faces.reserve(3 * nf);
	uint64_t first = 0;
	for (uint64_t f = 0; f < nf; f++) {
		uint64_t k = 0;
		if (!is.num(k)) return fail(GP_ERR_INVALID, "off: face %llu unreadable", (unsigned long long) f);
		if (f == 0) first = k;
		if (k != first || (k != 3 && k != 4))
			return fail(GP_ERR_INVALID, "off: face %llu has %llu corners (triangles or quads only, not mixed)",
			            (unsigned long long) f, (unsigned long long) k);
		uint32_t c[4];
		for (uint64_t q = 0; q < k; q++)
			if (!is.num(c[q])) return fail(GP_ERR_INVALID, "off: face %llu index unreadable", (unsigned long long) f);
		faces.push_back(c[0]);
		faces.push_back(c[1]);
		faces.push_back(c[2]);
		if (k == 4) {
			faces.push_back(c[3]);
			faces.push_back(c[0]);
			faces.push_back(c[2]);
		}
	}
	return GP_OK;
}

static int read_obj(Text &is, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	uint64_t ln = 0;
	std::vector<int64_t> idx;
	std::string key;
	Text l;
	while (is.line(l)) {
		ln++;
		if (!l.word(key)) continue;
		if (key == "v") {
			double x, y, z;
			if (!(l.num(x) && l.num(y) && l.num(z))) return fail(GP_ERR_INVALID, "obj: line %llu: bad vertex", (unsigned long long) ln);
			xyz.push_back(x);
			xyz.push_back(y);
			xyz.push_back(z);
		} else if (key == "f") {
			idx.clear();
			for (;;) {
				// The vertex index: the token up to its first '/', a positive
				// decimal integer (an optional '+' sign) and nothing else.
				l.skip();
				if (l.p >= l.e) break;
				const char *t0 = l.p;
				while (l.p < l.e && !Text::space(*l.p)) l.p++;
				const char *slash = static_cast<const char *>(std::memchr(t0, '/', (size_t) (l.p - t0)));
				const char *b = t0, *e = slash ? slash : l.p;
				if (e - b > 1 && *b == '+') b++;
				long long i = 0;
				const auto r = std::from_chars(b, e, i);
				if (b == e || r.ec != std::errc() || r.ptr != e || i <= 0) {
					const std::string tok(t0, l.p);
					return fail(GP_ERR_INVALID, "obj: line %llu: bad face index '%s'", (unsigned long long) ln, tok.c_str());
				}
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

// PLY scalar types, decoded little-endian into a double (indices and counts
// as unsigned where the type is).
enum PlyType { PLY_NONE, PLY_I8, PLY_U8, PLY_I16, PLY_U16, PLY_I32, PLY_U32, PLY_F32, PLY_F64 };

static PlyType ply_type(const std::string &t)
{
	if (t == "char" || t == "int8") return PLY_I8;
	if (t == "uchar" || t == "uint8") return PLY_U8;
	if (t == "short" || t == "int16") return PLY_I16;
	if (t == "ushort" || t == "uint16") return PLY_U16;
	if (t == "int" || t == "int32") return PLY_I32;
	if (t == "uint" || t == "uint32") return PLY_U32;
	if (t == "float" || t == "float32") return PLY_F32;
	if (t == "double" || t == "float64") return PLY_F64;
	return PLY_NONE;
}

static size_t ply_size(PlyType t)
{
	switch (t) {
	case PLY_I8: case PLY_U8: return 1;
	case PLY_I16: case PLY_U16: return 2;
	case PLY_I32: case PLY_U32: case PLY_F32: return 4;
	case PLY_F64: return 8;
	default: return 0;
	}
}

static double ply_value(const char *p, PlyType t)
{
	switch (t) {
	case PLY_I8: return (double) (int8_t) p[0];
	case PLY_U8: return (double) (uint8_t) p[0];
	case PLY_I16: { int16_t x; std::memcpy(&x, p, 2); return x; }
	case PLY_U16: { uint16_t x; std::memcpy(&x, p, 2); return x; }
	case PLY_I32: { int32_t x; std::memcpy(&x, p, 4); return x; }
	case PLY_U32: { uint32_t x; std::memcpy(&x, p, 4); return x; }
	case PLY_F32: { float x; std::memcpy(&x, p, 4); return x; }
	default: { double x; std::memcpy(&x, p, 8); return x; }
	}
}

static int read_ply(Text &is, std::vector<double> &xyz, std::vector<uint32_t> &faces)
{
	std::string format, element;
	uint64_t nv = 0, nf = 0;
	std::vector<PlyType> vtypes;              // vertex property types in order
	PlyType cnt_t = PLY_NONE, idx_t = PLY_NONE;   // face list types
	Text l;
	if (!is.line(l) || std::string(l.p, l.e).rfind("ply", 0) != 0) return fail(GP_ERR_INVALID, "ply: missing header");
	bool end = false;
	while (is.line(l)) {
		std::istringstream ss(std::string(l.p, l.e));
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
			if (!ply_type(t)) return fail(GP_ERR_INVALID, "ply: vertex property type '%s'", t.c_str());
			vtypes.push_back(ply_type(t));
		} else if (w == "property" && element == "face") {
			std::string t, c, i;
			ss >> t;
			if (t == "list") {
				ss >> c >> i;
				cnt_t = ply_type(c);
				idx_t = ply_type(i);
			}
		}
	}
	if (!end) return fail(GP_ERR_INVALID, "ply: no end_header");
	const auto fl = [](PlyType t) { return t == PLY_F32 || t == PLY_F64; };
	if (vtypes.size() < 3 || !fl(vtypes[0])) return fail(GP_ERR_INVALID, "ply: x y z must be the first vertex properties");
	if (nv >= SLOT_WRAP || nf > (1ull << 31)) return fail(GP_ERR_CAPACITY, "ply: counts too large");
	if (nf && (!cnt_t || !idx_t)) return fail(GP_ERR_INVALID, "ply: face list types");
	xyz.resize(3 * nv);
	faces.clear();
	faces.reserve(3 * nf);
	if (format == "ascii") {
		// One vertex or face per line; what follows the values read is ignored.
		for (uint64_t v = 0; v < nv; v++) {
			if (!is.line(l)) return fail(GP_ERR_INVALID, "ply: vertex %llu missing", (unsigned long long) v);
			if (!(l.num(xyz[3 * v]) && l.num(xyz[3 * v + 1]) && l.num(xyz[3 * v + 2])))
				return fail(GP_ERR_INVALID, "ply: vertex %llu unreadable", (unsigned long long) v);
		}
		for (uint64_t f = 0; f < nf; f++) {
			if (!is.line(l)) return fail(GP_ERR_INVALID, "ply: face %llu missing", (unsigned long long) f);
			uint64_t k = 0;
			uint32_t c[3];
			if (!l.num(k) || k != 3 || !(l.num(c[0]) && l.num(c[1]) && l.num(c[2])))
				return fail(GP_ERR_INVALID, "ply: face %llu is not a triangle", (unsigned long long) f);
			faces.push_back(c[0]);
			faces.push_back(c[1]);
			faces.push_back(c[2]);
		}
	} else if (format == "binary_little_endian") {
		size_t vbytes = 0;
		for (PlyType t : vtypes) vbytes += ply_size(t);
		const size_t s0 = ply_size(vtypes[0]), s1 = ply_size(vtypes[1]);
		const char *p = is.p;
		for (uint64_t v = 0; v < nv; v++, p += vbytes) {
			if ((size_t) (is.e - p) < vbytes) return fail(GP_ERR_INVALID, "ply: vertex %llu truncated", (unsigned long long) v);
			xyz[3 * v] = ply_value(p, vtypes[0]);
			xyz[3 * v + 1] = ply_value(p + s0, vtypes[1]);
			xyz[3 * v + 2] = ply_value(p + s0 + s1, vtypes[2]);
		}
		const size_t cs = ply_size(cnt_t), is_ = ply_size(idx_t);
		faces.resize(3 * nf);
		for (uint64_t f = 0; f < nf; f++) {
			if ((size_t) (is.e - p) < cs) return fail(GP_ERR_INVALID, "ply: face %llu truncated", (unsigned long long) f);
			const double k = ply_value(p, cnt_t);
			p += cs;
			if (k != 3) return fail(GP_ERR_INVALID, "ply: face %llu is not a triangle", (unsigned long long) f);
			if ((size_t) (is.e - p) < 3 * is_) return fail(GP_ERR_INVALID, "ply: face %llu truncated", (unsigned long long) f);
			for (int q = 0; q < 3; q++, p += is_) {
				const double i = ply_value(p, idx_t);
				if (i < 0) return fail(GP_ERR_INVALID, "ply: face %llu negative index", (unsigned long long) f);
				faces[3 * f + q] = (uint32_t) i;
			}
		}
		is.p = p;
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
	const bool off = ends_with(p, ".off"), obj = ends_with(p, ".obj"), ply = ends_with(p, ".ply");
	if (!off && !obj && !ply) return fail(GP_ERR_INVALID, "mesh file: unknown extension of '%s' (.off, .obj, .ply)", path);
	// The whole file in one read.
	std::string buf;
	is.seekg(0, std::ios::end);
	const std::streamoff size = is.tellg();
	if (size < 0) return fail(GP_ERR_INVALID, "mesh file: cannot size '%s'", path);
	buf.resize((size_t) size);
	is.seekg(0, std::ios::beg);
	if (size && !is.read(&buf[0], size)) return fail(GP_ERR_INVALID, "mesh file: cannot read '%s'", path);
	Text t{ buf.data(), buf.data() + buf.size() };
	return off ? read_off(t, xyz, faces) : obj ? read_obj(t, xyz, faces) : read_ply(t, xyz, faces);
}

}  // namespace gp
