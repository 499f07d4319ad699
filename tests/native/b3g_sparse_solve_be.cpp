/*
 * b3SparseSolveBE (sparse/sparse_force_solver.h:97) on the GPU: the translation
 * unit an integration puts in place of the reference's
 * sparse/sparse_force_solver.cpp. Every caller -- b3ForceSolver::Solve
 * (force_solver.cpp:88-180) and with it b3Body::Step -- keeps its code; the
 * force model (b3ForceModel: particles, forces, contacts) still runs on the
 * host, into the reference's own b3SparseMat33, and each row is handed over
 * head to tail with b3g_matrix_set_row. The linearisation, the pre-filtered
 * PCG and the Newton update run in libb3g (b3g_solve_be).
 *
 * oracle/Makefile builds it with the reference's other sources (target
 * ref_be) into a test library; tests/test_gpu_reference_be.py steps the
 * reference's bodies through it against the reference's own trajectories.
 */
#include <bounce_softbody/sparse/dense_vec3.h>
#include <bounce_softbody/sparse/diag_mat33.h>
#include <bounce_softbody/sparse/sparse_force_solver.h>
#include <bounce_softbody/sparse/sparse_mat33.h>

#include "b3g.h"

#include <cstring>
#include <vector>

static_assert(sizeof(scalar) == sizeof(float), "b3g_solve_be computes in FP32: build the reference with scalar = float");

namespace {

b3g_be g_solver = nullptr;
int g_reduction = B3G_REDUCE_FP64;
b3g_status g_last = B3G_OK;

// b3SparseForceModel::ApplyForces on the host, for the device solve.
b3g_status ApplyOnHost(void* user, const b3g_force_data* d) {
	b3SparseForceModel* model = static_cast<b3SparseForceModel*>(user);
	const uint32 n = d->n;
	b3DenseVec3 x(n), v(n), f(n);
	for (uint32 i = 0; i < n; ++i) {
		x[i].Set(d->x[3 * i], d->x[3 * i + 1], d->x[3 * i + 2]);
		v[i].Set(d->v[3 * i], d->v[3 * i + 1], d->v[3 * i + 2]);
	}
	f.SetZero();
	b3SparseMat33 dfdx(n), dfdv(n);
	b3SparseForceSolverData data;
	data.h = d->h;
	data.inv_h = d->inv_h;
	data.x = &x;
	data.v = &v;
	data.f = &f;
	data.dfdx = &dfdx;
	data.dfdv = &dfdv;
	model->ApplyForces(&data);
	for (uint32 i = 0; i < n; ++i) {
		d->f[3 * i] = f[i].x;
		d->f[3 * i + 1] = f[i].y;
		d->f[3 * i + 2] = f[i].z;
	}
	std::vector<uint32_t> cols;
	std::vector<float> blocks;
	const b3SparseMat33* src[2] = {&dfdx, &dfdv};
	b3g_matrix dst[2] = {d->dfdx, d->dfdv};
	for (int m = 0; m < 2; ++m)
		for (uint32 i = 0; i < n; ++i) {
			cols.clear();
			blocks.clear();
			for (const b3RowEntry* e = src[m]->rows[i].head; e; e = e->next) {   // head to tail
				cols.push_back(e->column);
				const float* b = &e->value.x.x;
				blocks.insert(blocks.end(), b, b + 9);
			}
			if (cols.empty()) continue;
			const b3g_status s = b3g_matrix_set_row(dst[m], i, (uint32_t)cols.size(), cols.data(), blocks.data());
			if (s != B3G_OK) return s;
		}
	return B3G_OK;
}

} // namespace

// The reduction mode of the device solve (B3G_REDUCE_SEQUENTIAL: the reference's results bit for bit) and the
// status of the last solve; test hooks, named for the harness's export map.
extern "C" void refh_be_set_reduction(int mode) {
	g_reduction = mode;
	if (g_solver) b3g_be_set_reduction(g_solver, mode);
}

extern "C" int refh_be_last_status(void) { return (int)g_last; }

void b3SparseSolveBE(b3SolveBEOutput* output, const b3SolveBEInput* input) {
	const uint32 n = input->dofCount;
	if (!g_solver) {
		g_last = b3g_be_create(&g_solver, 0);
		if (g_last != B3G_OK) {
			b3Log("b3g: %s\n", b3g_last_error());
			g_solver = nullptr;
			return;
		}
		b3g_be_set_reduction(g_solver, g_reduction);
	}
	b3g_be_input in;
	b3g_be_input_defaults(&in);
	in.h = input->h;
	in.inv_h = input->inv_h;
	in.force_model = ApplyOnHost;
	in.user = input->forceModel;
	in.n = n;
	if (n) {
		in.x0 = &input->x0->v[0].x;   // b3DenseVec3 / b3DiagMat33 hold contiguous b3Vec3 / b3Mat33 arrays
		in.v0 = &input->v0->v[0].x;
		in.fe = &input->fe->v[0].x;
		in.M = &input->M->v[0].x.x;
		in.y = &input->y->v[0].x;
		in.S = &input->S->v[0].x.x;
		in.z = &input->z->v[0].x;
	}
	in.max_iterations = input->maxIterations;
	in.tolerance = input->tolerance;
	in.max_sub_iterations = input->maxSubIterations;
	in.sub_tolerance = input->subTolerance;
	b3g_be_output out;
	std::memset(&out, 0, sizeof(out));
	out.x = n ? &output->x->v[0].x : nullptr;
	out.v = n ? &output->v->v[0].x : nullptr;
	out.min_sub_iterations = output->minSubIterations;
	out.max_sub_iterations = output->maxSubIterations;
	g_last = b3g_solve_be(g_solver, &out, &in);
	if (g_last != B3G_OK) b3Log("b3g: %s\n", b3g_last_error());
	if (g_last == B3G_OK || g_last == B3G_NON_POSITIVE_DIAGONAL || g_last == B3G_NON_FINITE) {
		// results written as the reference's release build would write them
		output->iterations = out.iterations;
		output->error = out.error;
		output->minSubIterations = out.min_sub_iterations;
		output->maxSubIterations = out.max_sub_iterations;
	} else {
		*output->x = *input->x0;   // no device result: the state does not move
		*output->v = *input->v0;
		output->iterations = 0;
		output->error = scalar(0);
	}
}
