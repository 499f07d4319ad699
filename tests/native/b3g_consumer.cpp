// A C++ program that uses libb3g the way a patched b3Body::Step would
// (include/b3g.h only; linked against bounce_softbody_amd/libb3g.so): it
// builds a body from a description, steps it on a schedule and writes the
// state after every step. The test that drives it (tests/test_native_consumer.py)
// writes the description, compares the states with a fixture written by the
// reference, and checks the error paths.
//
//   b3g_consumer <scene.bin> <out.bin>
//
// scene.bin: the description's arrays in b3g_body_desc order, each as
// (uint32 element size, uint64 count, data); then the gravity (3 floats); then
// the schedule: uint64 steps, per step (float dt, uint32 iterations, uint32
// sub_iterations, float tolerance, float sub_tolerance). World fixtures are
// limited to analytic shapes (no SDF voxels or meshes).
// out.bin: per step, the positions then the velocities (3 floats per particle).
// Exit status 0 on success; on a library error, the status and
// b3g_last_error() go to stderr and the status is the exit code.
#include "b3g.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

struct Reader {
	FILE* f;
	bool ok = true;
	template <class T> std::vector<T> array() {
		uint32_t size = 0;
		uint64_t count = 0;
		ok = ok && std::fread(&size, 4, 1, f) == 1 && std::fread(&count, 8, 1, f) == 1 && size == sizeof(T);
		std::vector<T> v(ok ? count : 0);
		if (ok && count) ok = std::fread(v.data(), sizeof(T), count, f) == count;
		return v;
	}
	template <class T> T value() {
		T v{};
		ok = ok && std::fread(&v, sizeof(T), 1, f) == 1;
		return v;
	}
};

template <class T> const T* ptr(const std::vector<T>& v) { return v.empty() ? nullptr : v.data(); }

int fail(const char* where, b3g_status s) {
	std::fprintf(stderr, "%s: %s (%d): %s\n", where, b3g_status_string(s), (int)s, b3g_last_error());
	return (int)s;
}

} // namespace

int main(int argc, char** argv) {
	if (argc != 3) {
		std::fprintf(stderr, "usage: %s scene.bin out.bin\n", argv[0]);
		return 64;
	}
	FILE* in = std::fopen(argv[1], "rb");
	if (!in) { std::perror(argv[1]); return 66; }
	Reader r{in};
	auto p_type = r.array<uint8_t>();
	auto p_pos = r.array<float>(), p_vel = r.array<float>(), p_damping = r.array<float>();
	auto s_particle = r.array<uint32_t>();
	auto s_radius = r.array<float>(), s_friction = r.array<float>();
	auto tf_p = r.array<uint32_t>();
	auto tf_v = r.array<float>(), tf_density = r.array<float>(), tf_radius = r.array<float>(), tf_friction = r.array<float>();
	auto ttf_p = r.array<uint32_t>();
	auto ttf_v = r.array<float>(), ttf_density = r.array<float>(), ttf_radius = r.array<float>(),
	     ttf_friction = r.array<float>();
	auto f_type = r.array<uint32_t>(), f_index = r.array<uint32_t>();
	auto tet_p = r.array<uint32_t>();
	auto tet_v = r.array<float>(), tet_param = r.array<float>();
	auto tri_p = r.array<uint32_t>();
	auto tri_v = r.array<float>(), tri_param = r.array<float>();
	auto st_p = r.array<uint32_t>();
	auto st_param = r.array<float>();
	auto sh_p = r.array<uint32_t>();
	auto sh_param = r.array<float>();
	auto sp_p = r.array<uint32_t>();
	auto sp_param = r.array<float>();
	auto mo_p = r.array<uint32_t>();
	auto mo_param = r.array<float>();
	auto w_type = r.array<uint32_t>();
	auto w_param = r.array<float>(), w_friction = r.array<float>();
	float gravity[3];
	for (float& g : gravity) g = r.value<float>();
	const uint64_t steps = r.value<uint64_t>();
	struct Row { float dt; uint32_t it, sub; float tol, stol; };
	std::vector<Row> schedule(r.ok ? steps : 0);
	for (Row& row : schedule) {
		row.dt = r.value<float>();
		row.it = r.value<uint32_t>();
		row.sub = r.value<uint32_t>();
		row.tol = r.value<float>();
		row.stol = r.value<float>();
	}
	std::fclose(in);
	if (!r.ok) { std::fprintf(stderr, "%s: malformed scene file\n", argv[1]); return 65; }

	b3g_body_desc d;
	std::memset(&d, 0, sizeof(d));
	d.n_particles = (uint32_t)p_type.size();
	d.p_type = ptr(p_type); d.p_pos = ptr(p_pos); d.p_vel = ptr(p_vel); d.p_damping = ptr(p_damping);
	d.n_spheres = (uint32_t)s_particle.size();
	d.s_particle = ptr(s_particle); d.s_radius = ptr(s_radius); d.s_friction = ptr(s_friction);
	d.n_trifix = (uint32_t)tf_density.size();
	d.tf_p = ptr(tf_p); d.tf_v = ptr(tf_v); d.tf_density = ptr(tf_density); d.tf_radius = ptr(tf_radius);
	d.tf_friction = ptr(tf_friction);
	d.n_tetfix = (uint32_t)ttf_density.size();
	d.ttf_p = ptr(ttf_p); d.ttf_v = ptr(ttf_v); d.ttf_density = ptr(ttf_density); d.ttf_radius = ptr(ttf_radius);
	d.ttf_friction = ptr(ttf_friction);
	d.n_forces = (uint32_t)f_type.size();
	d.f_type = ptr(f_type); d.f_index = ptr(f_index);
	d.n_tet = (uint32_t)(tet_param.size() / 3);
	d.tet_p = ptr(tet_p); d.tet_v = ptr(tet_v); d.tet_param = ptr(tet_param);
	d.n_tri = (uint32_t)(tri_param.size() / 6);
	d.tri_p = ptr(tri_p); d.tri_v = ptr(tri_v); d.tri_param = ptr(tri_param);
	d.n_stretch = (uint32_t)(st_param.size() / 12);
	d.st_p = ptr(st_p); d.st_param = ptr(st_param);
	d.n_shear = (uint32_t)(sh_param.size() / 8);
	d.sh_p = ptr(sh_p); d.sh_param = ptr(sh_param);
	d.n_spring = (uint32_t)(sp_param.size() / 3);
	d.sp_p = ptr(sp_p); d.sp_param = ptr(sp_param);
	d.n_mouse = (uint32_t)(mo_param.size() / 6);
	d.mo_p = ptr(mo_p); d.mo_param = ptr(mo_param);
	d.n_world = (uint32_t)w_type.size();
	d.w_type = ptr(w_type); d.w_param = ptr(w_param); d.w_friction = ptr(w_friction);
	std::vector<const float*> voxels(d.n_world, nullptr);
	d.w_voxels = voxels.empty() ? nullptr : voxels.data();
	std::memcpy(d.gravity, gravity, sizeof(gravity));

	int count = 0;
	b3g_status s = b3g_device_count(&count);
	if (s != B3G_OK) return fail("b3g_device_count", s);
	if (count < 1) { std::fprintf(stderr, "no device\n"); return 69; }
	b3g_body body = nullptr;
	if ((s = b3g_body_create(&body, 0)) != B3G_OK) return fail("b3g_body_create", s);
	// Every status is kept in a variable and checked: none of these calls sits inside an assert.
	s = b3g_body_build(body, &d);
	if (s != B3G_OK) { b3g_body_destroy(body); return fail("b3g_body_build", s); }
	// the reference's own summation order, so the states equal the reference's bit for bit
	s = b3g_set_reduction(body, B3G_REDUCE_SEQUENTIAL);
	if (s != B3G_OK) { b3g_body_destroy(body); return fail("b3g_set_reduction", s); }
	FILE* out = std::fopen(argv[2], "wb");
	if (!out) { std::perror(argv[2]); b3g_body_destroy(body); return 73; }
	std::vector<float> x(3 * (size_t)d.n_particles), v(3 * (size_t)d.n_particles);
	int rc = 0;
	for (const Row& row : schedule) {
		s = b3g_step(body, row.dt, row.it, row.sub, row.tol, row.stol);
		if (s != B3G_OK) { rc = fail("b3g_step", s); break; }
		s = b3g_get_state(body, x.data(), v.data());
		if (s != B3G_OK) { rc = fail("b3g_get_state", s); break; }
		if (std::fwrite(x.data(), 4, x.size(), out) != x.size() || std::fwrite(v.data(), 4, v.size(), out) != v.size()) {
			std::perror(argv[2]);
			rc = 74;
			break;
		}
	}
	std::fclose(out);
	b3g_body_destroy(body);
	return rc;
}
