// Box persistence in the reference's particle file layout.
//
// particles_save / particles_load (src/particles.cpp:1182-1227) write, in
// order: the particle count (int64), x, y, z (fixed32 each), the velocities
// (float[3] per particle), the rungs (one byte each), then the active-range
// stack: its size (int32), rung_begin (int32) and the saved starts (int32
// each).  The optional group and tracer arrays (do_groups, do_tracers) are
// not part of a box here.  The checkpoint of write_checkpoint
// (src/driver.cpp:1143-1172) puts driver_params in front; ct_box_save can
// append to such a file and ct_box_load can start at an offset.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "ct_host.hpp"

using namespace ct;

namespace {

struct File {
	FILE* f = nullptr;
	~File() {
		if (f) std::fclose(f);
	}
};

}  // namespace

int ct_box_save(ct_box_t b, const char* path, int append) {
	if (!b || !path) return fail(CT_E_ARG, "ct_box_save: bad arguments");
	const int64_t n = b->hp.n;
	std::vector<uint32_t> x(n), y(n), z(n);
	std::vector<float> vel(3 * n);
	std::vector<int8_t> rung(n);
	int rc = ct_box_read(b, x.data(), y.data(), z.data(), vel.data(), rung.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
	if (rc) return rc;
	if (b->rung_begin > INT32_MAX || b->rung_stack.size() > size_t(INT32_MAX))
		return fail(CT_E_ARG, "ct_box_save: range start does not fit the file's int32");
	File fp;
	fp.f = std::fopen(path, append ? "ab" : "wb");
	if (!fp.f) return fail(CT_E_ARG, std::string("ct_box_save: cannot open ") + path);
	bool ok = std::fwrite(&n, sizeof(int64_t), 1, fp.f) == 1;
	ok = ok && std::fwrite(x.data(), 4, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fwrite(y.data(), 4, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fwrite(z.data(), 4, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fwrite(vel.data(), 12, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fwrite(rung.data(), 1, size_t(n), fp.f) == size_t(n);
	const int32_t ns = int32_t(b->rung_stack.size());
	const int32_t rb = int32_t(b->rung_begin);
	ok = ok && std::fwrite(&ns, 4, 1, fp.f) == 1;
	ok = ok && std::fwrite(&rb, 4, 1, fp.f) == 1;
	for (int64_t v : b->rung_stack) {
		const int32_t w = int32_t(v);
		ok = ok && std::fwrite(&w, 4, 1, fp.f) == 1;
	}
	ok = ok && std::fflush(fp.f) == 0;
	if (!ok) return fail(CT_E_ARG, std::string("ct_box_save: write failed on ") + path);
	return CT_OK;
}

int ct_box_load(const char* path, int64_t offset, ct_box_t* out, int64_t* end_offset) {
	int rc = check_init();
	if (rc) return rc;
	if (!path || !out || offset < 0) return fail(CT_E_ARG, "ct_box_load: bad arguments");
	File fp;
	fp.f = std::fopen(path, "rb");
	if (!fp.f) return fail(CT_E_ARG, std::string("ct_box_load: cannot open ") + path);
	if (std::fseek(fp.f, long(offset), SEEK_SET) != 0) return fail(CT_E_ARG, "ct_box_load: bad offset");
	int64_t n = 0;
	if (std::fread(&n, sizeof(int64_t), 1, fp.f) != 1 || n < 0 || n > INT32_MAX)
		return fail(CT_E_ARG, std::string("ct_box_load: bad particle count in ") + path);
	std::vector<uint32_t> x(n), y(n), z(n);
	std::vector<float> vel(3 * n);
	std::vector<int8_t> rung(n);
	bool ok = std::fread(x.data(), 4, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fread(y.data(), 4, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fread(z.data(), 4, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fread(vel.data(), 12, size_t(n), fp.f) == size_t(n);
	ok = ok && std::fread(rung.data(), 1, size_t(n), fp.f) == size_t(n);
	int32_t ns = 0, rb = 0;
	ok = ok && std::fread(&ns, 4, 1, fp.f) == 1 && std::fread(&rb, 4, 1, fp.f) == 1;
	if (!ok || ns < 0 || rb < 0 || rb > n) return fail(CT_E_ARG, std::string("ct_box_load: truncated or corrupt file ") + path);
	std::vector<int32_t> stack(ns);
	if (ns > 0 && std::fread(stack.data(), 4, size_t(ns), fp.f) != size_t(ns))
		return fail(CT_E_ARG, std::string("ct_box_load: truncated range stack in ") + path);
	for (int32_t v : stack)
		if (v < 0 || v > n) return fail(CT_E_ARG, std::string("ct_box_load: corrupt range stack in ") + path);
	const long end = std::ftell(fp.f);
	ct_box_t b = nullptr;
	if ((rc = ct_box_create(n, x.data(), y.data(), z.data(), vel.data(), rung.data(), &b))) return rc;
	b->rung_begin = rb;
	b->rung_stack.assign(stack.begin(), stack.end());
	*out = b;
	if (end_offset) *end_offset = int64_t(end);
	return CT_OK;
}
