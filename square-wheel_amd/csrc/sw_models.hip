// Triangle-model preparation on the GPU (triangle_models_rendering.rs:7-200, 324-391, 702-731,
// light.rs:296-340): one block per model instance animates, transforms and lights the mesh's
// vertices, then rejects back faces and sorts the front-facing triangles back to front.
//
// Float operations keep the reference's order and fusion (cgmath 0.18 f32 arithmetic: matrix x
// vector is the column sum c0*x + c1*y + c2*z (+ c3*w), left to right, unfused; the file is built
// with -ffp-contract=off), so the results are bit-exact to the oracle's restatement.
#include "sw_internal.h"
#include "sw_device.h"

using namespace swd;

namespace
{

#define MODEL_BLOCK 1024u
#define MODEL_SORT_LDS 4096u // sort keys kept in LDS up to this many (power of two) triangles

struct F3
{
	float x, y, z;
};
SW_DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
SW_DEV F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
SW_DEV float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SW_DEV F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// Column-major matrices: m[col * N + row].
struct M3
{
	float m[9];
};
struct M4
{
	float m[16];
};

// Matrix4 * (v, w), truncated to xyz.
SW_DEV F3 m4_transform(const float* m, F3 v, float w)
{
	return f3(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
		m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w);
}
SW_DEV F3 m3_mul(const float* m, F3 v)
{
	return f3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z, m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
// a * b (columns of the product: a * b.col_j)
SW_DEV void m4_mul(const float* a, const float* b, float* out)
{
	for(int c = 0; c < 4; ++c)
		for(int r = 0; r < 4; ++r)
			out[c * 4 + r] = a[r] * b[c * 4] + a[4 + r] * b[c * 4 + 1] + a[8 + r] * b[c * 4 + 2] + a[12 + r] * b[c * 4 + 3];
}
SW_DEV void m3_mul_m3(const float* a, const float* b, float* out)
{
	for(int c = 0; c < 3; ++c)
		for(int r = 0; r < 3; ++r)
			out[c * 3 + r] = a[r] * b[c * 3] + a[3 + r] * b[c * 3 + 1] + a[6 + r] * b[c * 3 + 2];
}

// get_normals_matrix (triangle_models_rendering.rs:722-731): the 3x3 axes of the model matrix,
// transposed and inverted (cgmath 0.18 Matrix3::invert: cofactor columns c1 x c2, c2 x c0, c0 x c1
// divided by the determinant, transposed; identity when the determinant is 0).
SW_DEV void normals_matrix(const float* m4, float* out)
{
	// A = transpose(axes): column k of A is row k of the axes, A[k][j] = m4 column j, component k
	const F3 c0 = f3(m4[0], m4[4], m4[8]), c1 = f3(m4[1], m4[5], m4[9]), c2 = f3(m4[2], m4[6], m4[10]);
	const float det = c0.x * (c1.y * c2.z - c2.y * c1.z) - c1.x * (c0.y * c2.z - c2.y * c0.z) + c2.x * (c0.y * c1.z - c1.y * c0.z);
	if(det == 0.0f)
	{
		for(int i = 0; i < 9; ++i)
			out[i] = (i % 4 == 0) ? 1.0f : 0.0f;
		return;
	}
	const F3 r0 = cross(c1, c2), r1 = cross(c2, c0), r2 = cross(c0, c1);
	const F3 r[3] = {f3(r0.x / det, r0.y / det, r0.z / det), f3(r1.x / det, r1.y / det, r1.z / det),
		f3(r2.x / det, r2.y / det, r2.z / det)};
	// transpose of from_cols(r0, r1, r2): column j = (r0[j], r1[j], r2[j])
	for(int j = 0; j < 3; ++j)
		for(int k = 0; k < 3; ++k)
			out[j * 3 + k] = (&r[k].x)[j];
}

// get_vertex_light (triangle_models_rendering.rs:702-720) with get_light_cube_light (light.rs:296-340).
SW_DEV void vertex_light(const sw_model_instance& in, F3 n, float* out)
{
	const float s = inv_sqrt_fast(fmax_r(dot(n, n), 0.00000001f));
	const F3 nn = f3(n.x * s, n.y * s, n.z * s);
	float t[3] = {0.0f, 0.0f, 0.0f};
	const float comp[3] = {nn.x, nn.y, nn.z};
	for(int axis = 0; axis < 3; ++axis)
	{
		const bool neg = comp[axis] <= 0.0f;
		const float k = neg ? -comp[axis] : comp[axis];
		const float* side = in.light_cube[axis * 2 + (neg ? 0 : 1)];
		for(int i = 0; i < 3; ++i)
			t[i] += side[i] * k;
	}
	for(int d = 0; d < 2; ++d)
	{
		const F3 v = f3(in.light_vector_scaled[d][0], in.light_vector_scaled[d][1], in.light_vector_scaled[d][2]);
		const float dd = fmax_r(dot(nn, v), 0.0f);
		for(int i = 0; i < 3; ++i)
			t[i] += in.light_color[d][i] * dd;
	}
	out[0] = t[0];
	out[1] = t[1];
	out[2] = t[2];
}

SW_DEV void store_vertex(sw_model_vertex* dst, F3 pos, float tc0, float tc1, const float* light)
{
	float4* d = reinterpret_cast<float4*>(dst);
	d[0] = make_float4(pos.x, pos.y, pos.z, tc0);
	d[1] = make_float4(tc1, light[0], light[1], light[2]);
}

SW_DEV F3 load_f3(const float* p) { return f3(p[0], p[1], p[2]); }

// Ascending order of (depth descending, triangle index ascending): the depth's bits mapped to an
// unsigned order and inverted, the index below.
SW_DEV uint64_t sort_key(float z, uint32_t t)
{
	uint32_t u = __float_as_uint(z == 0.0f ? 0.0f : z);
	u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
	return (uint64_t(~u) << 32) | t;
}

struct ModelLaunch
{
	const sw_model_mesh* meshes;
	const sw_model_instance* instances;
	const uint64_t* sort_offsets; // per instance: its region of `scratch` (sorts above MODEL_SORT_LDS)
	uint64_t* scratch;
	sw_model_vertex* out_vertices;
	uint16_t* out_triangles;
	uint32_t* out_counts;
	uint32_t* host_err; // pinned word set on an out-of-range vertex or bone index
};

__global__ __launch_bounds__(MODEL_BLOCK) void k_prepare_models(ModelLaunch L)
{
	__shared__ float bone_m[SW_MODEL_MAX_BONES * 16]; // model-view * bone matrix, / 255
	__shared__ float bone_n[SW_MODEL_MAX_BONES * 9]; // normals matrix * bone normals matrix, / 255
	__shared__ uint8_t bone_done[SW_MODEL_MAX_BONES];
	__shared__ uint64_t keys_lds[MODEL_SORT_LDS];
	__shared__ uint32_t visible;

	const sw_model_instance& in = L.instances[blockIdx.x];
	const sw_model_mesh M = L.meshes[in.mesh];
	const uint32_t tid = threadIdx.x;
	bool bad = false;

	const float lerp0 = fmin_r(fmax_r(in.lerp, 0.0f), 1.0f);
	const float lerp1 = 1.0f - lerp0;
	const bool perform_lerp = lerp0 > 0.01f && lerp0 < 0.99f;
	const uint32_t frame_single = lerp0 > lerp1 ? in.frames[0] : in.frames[1];
	float nm[9];
	normals_matrix(in.model_matrix, nm);
	sw_model_vertex* out_v = L.out_vertices + in.first_vertex;

	const bool skeleton = M.kind == SW_MODEL_SKELETON_ANIMATED && M.num_bones > 0;
	if(skeleton)
	{
		// bone matrices in hierarchy order (:108-150), in rounds: a bone is computed (by one thread,
		// in the reference's operand order) in the round after its parent was. bone_done: 0 open,
		// 1 computed in an earlier round, 2 computed in this round. A parent index at or above
		// the bone's own (the reference requires parents first) counts as no parent and is an error.
		const uint32_t nb = M.num_bones;
		for(uint32_t b = tid; b < nb; b += MODEL_BLOCK)
			bone_done[b] = 0;
		__syncthreads();
		for(uint32_t round = 0; round < nb; ++round)
		{
			int open = 0;
			for(uint32_t b = tid; b < nb; b += MODEL_BLOCK)
			{
				if(bone_done[b])
					continue;
				uint32_t parent = M.bone_parents[b];
				if(parent < nb && parent >= b)
				{
					bad = true;
					parent = nb;
				}
				if(parent < nb && bone_done[parent] != 1)
				{
					open = 1;
					continue;
				}
				float local[16];
				if(perform_lerp)
				{
					const float* f0 = M.frame_bones + (uint64_t(in.frames[0]) * nb + b) * 16;
					const float* f1 = M.frame_bones + (uint64_t(in.frames[1]) * nb + b) * 16;
					for(int i = 0; i < 16; ++i)
						local[i] = f0[i] * lerp0 + f1[i] * lerp1;
				}
				else
				{
					const float* f = M.frame_bones + (uint64_t(frame_single) * nb + b) * 16;
					for(int i = 0; i < 16; ++i)
						local[i] = f[i];
				}
				if(parent < nb)
				{
					float pm[16], r[16];
					for(int i = 0; i < 16; ++i)
						pm[i] = bone_m[parent * 16 + i];
					m4_mul(pm, local, r);
					for(int i = 0; i < 16; ++i)
						bone_m[b * 16 + i] = r[i];
				}
				else
					for(int i = 0; i < 16; ++i)
						bone_m[b * 16 + i] = local[i];
				bone_done[b] = 2;
			}
			const int any_open = __syncthreads_or(open);
			for(uint32_t b = tid; b < nb; b += MODEL_BLOCK)
				if(bone_done[b] == 2)
					bone_done[b] = 1;
			__syncthreads();
			if(!any_open)
				break;
		}
		// (:152-165) weight-scaled normals and model-view products
		const float ws = 1.0f / 255.0f;
		float nmws[9], mvws[16];
		for(int i = 0; i < 9; ++i)
			nmws[i] = nm[i] * ws;
		for(int i = 0; i < 16; ++i)
			mvws[i] = in.model_view_matrix[i] * ws;
		for(uint32_t b = tid; b < nb; b += MODEL_BLOCK)
		{
			float bm[16], bn[9], r9[9], r16[16];
			for(int i = 0; i < 16; ++i)
				bm[i] = bone_m[b * 16 + i];
			normals_matrix(bm, bn);
			m3_mul_m3(nmws, bn, r9);
			m4_mul(mvws, bm, r16);
			for(int i = 0; i < 9; ++i)
				bone_n[b * 9 + i] = r9[i];
			for(int i = 0; i < 16; ++i)
				bone_m[b * 16 + i] = r16[i];
		}
		__syncthreads();
	}

	// vertices (:19-200)
	for(uint32_t v = tid; v < M.num_vertices; v += MODEL_BLOCK)
	{
		F3 pos, normal;
		float tc0, tc1;
		float light[3];
		if(M.kind == SW_MODEL_VERTEX_ANIMATED)
		{
			const float* tc = static_cast<const float*>(M.vertices) + uint64_t(v) * 2;
			tc0 = tc[0];
			tc1 = tc[1];
			if(perform_lerp)
			{
				const float* a = M.frame_vertices + (uint64_t(in.frames[0]) * M.num_vertices + v) * 6;
				const float* b = M.frame_vertices + (uint64_t(in.frames[1]) * M.num_vertices + v) * 6;
				pos = f3(a[0] * lerp0 + b[0] * lerp1, a[1] * lerp0 + b[1] * lerp1, a[2] * lerp0 + b[2] * lerp1);
				normal = f3(a[3] * lerp0 + b[3] * lerp1, a[4] * lerp0 + b[4] * lerp1, a[5] * lerp0 + b[5] * lerp1);
			}
			else
			{
				const float* a = M.frame_vertices + (uint64_t(frame_single) * M.num_vertices + v) * 6;
				pos = load_f3(a);
				normal = load_f3(a + 3);
			}
			vertex_light(in, m3_mul(nm, normal), light);
			pos = m4_transform(in.model_view_matrix, pos, 1.0f);
		}
		else if(M.kind == SW_MODEL_NON_ANIMATED || !skeleton)
		{
			// VertexNonAnimated {position, normal, tex_coord}; SkeletonAnimatedVertex without bones
			const float* p = static_cast<const float*>(M.vertices) + uint64_t(v) * (M.kind == SW_MODEL_NON_ANIMATED ? 8 : 10);
			if(M.kind == SW_MODEL_NON_ANIMATED)
			{
				pos = load_f3(p);
				normal = load_f3(p + 3);
				tc0 = p[6];
				tc1 = p[7];
			}
			else
			{
				tc0 = p[0];
				tc1 = p[1];
				pos = load_f3(p + 2);
				normal = load_f3(p + 5);
			}
			vertex_light(in, m3_mul(nm, normal), light);
			pos = m4_transform(in.model_view_matrix, pos, 1.0f);
		}
		else
		{
			// SkeletonAnimatedVertex (:167-198): the weighted sum of up to four bone matrices
			const float* p = static_cast<const float*>(M.vertices) + uint64_t(v) * 10;
			tc0 = p[0];
			tc1 = p[1];
			const F3 src_pos = load_f3(p + 2);
			normal = load_f3(p + 5);
			const uint32_t bw0 = reinterpret_cast<const uint32_t*>(p + 8)[0], bw1 = reinterpret_cast<const uint32_t*>(p + 8)[1];
			const uint32_t desc[4] = {bw0 & 0xFFFFu, bw0 >> 16, bw1 & 0xFFFFu, bw1 >> 16};
			float mat[16], nmat[9];
			for(int k = 0; k < 4; ++k)
			{
				uint32_t bi = desc[k] & 0xFFu;
				const uint32_t wi = desc[k] >> 8;
				if(k > 0 && wi == 0)
					continue;
				if(bi >= M.num_bones)
				{
					bad = true;
					bi = M.num_bones - 1;
				}
				const float w = float(wi);
				if(k == 0)
				{
					for(int i = 0; i < 16; ++i)
						mat[i] = bone_m[bi * 16 + i] * w;
					for(int i = 0; i < 9; ++i)
						nmat[i] = bone_n[bi * 9 + i] * w;
				}
				else
				{
					for(int i = 0; i < 16; ++i)
						mat[i] += bone_m[bi * 16 + i] * w;
					for(int i = 0; i < 9; ++i)
						nmat[i] += bone_n[bi * 9 + i] * w;
				}
			}
			vertex_light(in, m3_mul(nmat, normal), light);
			pos = m4_transform(mat, src_pos, 1.0f);
		}
		// tc: tex_coord * tc_scale + tc_shift (:37)
		store_vertex(out_v + v, pos, tc0 * in.tc_scale[0] + in.tc_shift[0], tc1 * in.tc_scale[1] + in.tc_shift[1], light);
	}
	if(tid == 0)
		visible = 0;
	__syncthreads();

	// back faces (:324-341): the front faces' sort keys (:343-391) are compacted to the front of
	// `keys` (in any order: the keys hold the triangle index, so the sort fixes the order)
	uint32_t n_cap = 1;
	while(n_cap < M.num_triangles)
		n_cap <<= 1;
	uint64_t* keys = n_cap <= MODEL_SORT_LDS ? keys_lds : L.scratch + L.sort_offsets[blockIdx.x];
	for(uint32_t t = tid; t < M.num_triangles; t += MODEL_BLOCK)
	{
		uint32_t idx[3];
		for(int k = 0; k < 3; ++k)
		{
			idx[k] = M.triangles[uint64_t(t) * 3 + k];
			if(idx[k] >= M.num_vertices)
			{
				bad = true;
				idx[k] = M.num_vertices - 1;
			}
		}
		const F3 p0 = load_f3(&out_v[idx[0]].pos[0]), p1 = load_f3(&out_v[idx[1]].pos[0]), p2 = load_f3(&out_v[idx[2]].pos[0]);
		// get_triangle_plane (:393-401): (v1 - v0) x (v2 - v1), dist = vec . v0
		const F3 vec = cross(sub(p1, p0), sub(p2, p1));
		if(dot(vec, p0) > 0.0f)
			keys[atomicAdd(&visible, 1u)] = sort_key(fmax_r(fmax_r(p0.z, p1.z), p2.z), t);
	}
	__syncthreads();
	const uint32_t count = visible;
	uint32_t n_pad = 1;
	while(n_pad < count)
		n_pad <<= 1;
	for(uint32_t t = count + tid; t < n_pad; t += MODEL_BLOCK)
		keys[t] = ~0ull;
	__syncthreads();

	// bitonic sort, ascending
	for(uint32_t k = 2; k <= n_pad; k <<= 1)
		for(uint32_t j = k >> 1; j > 0; j >>= 1)
		{
			for(uint32_t i = tid; i < n_pad; i += MODEL_BLOCK)
			{
				const uint32_t l = i ^ j;
				if(l > i)
				{
					const uint64_t a = keys[i], b = keys[l];
					if(((i & k) == 0) == (a > b))
					{
						keys[i] = b;
						keys[l] = a;
					}
				}
			}
			__syncthreads();
		}

	uint16_t* out_t = L.out_triangles + uint64_t(in.first_triangle) * 3;
	for(uint32_t i = tid; i < count; i += MODEL_BLOCK)
	{
		const uint32_t t = uint32_t(keys[i]);
		for(int k = 0; k < 3; ++k)
		{
			const uint32_t idx = M.triangles[uint64_t(t) * 3 + k];
			out_t[uint64_t(i) * 3 + k] = uint16_t(idx < M.num_vertices ? idx : M.num_vertices - 1);
		}
	}
	if(tid == 0)
		L.out_counts[blockIdx.x] = count;
	if(bad)
		*L.host_err = 1u;
}

// ---------------------------------------------------------------------------------------------
// draw_mesh's per-triangle part (partial_renderer.rs:3612-3716) over prepared models: clip by the
// run's 3D planes (clip_3d_model_polygon_by_plane, clipping.rs:259-313), project, clip by its 2D
// planes (clip_2d_model_polygon, :203-257), then the fan's triangles as fixed16 vertex data.
// ---------------------------------------------------------------------------------------------
#define MESH_MAX_POLY (3u + SW_MESH_MAX_PLANES3D + 4u)

struct MV
{
	float p[3]; // 3D position, or 2D position in p[0..1]
	float tc[2];
	float l[3];
};

// v0 * k1 - v1 * k0 with k0 = d0 / (d1 - d0), k1 = d1 / (d1 - d0) (get_model_line_*_intersection)
SW_DEV MV mv_intersect(const MV& v0, const MV& v1, float d0, float d1, int np)
{
	const float sum = d1 - d0;
	const float k0 = d0 / sum, k1 = d1 / sum;
	MV r;
	for(int i = 0; i < 3; ++i)
		r.p[i] = i < np ? v0.p[i] * k1 - v1.p[i] * k0 : 0.0f;
	for(int i = 0; i < 2; ++i)
		r.tc[i] = v0.tc[i] * k1 - v1.tc[i] * k0;
	for(int i = 0; i < 3; ++i)
		r.l[i] = v0.l[i] * k1 - v1.l[i] * k0;
	return r;
}

// One clip pass; `dist` of a vertex is dot(pos, plane) - offset, kept where >= 0 (the reference's
// dot > offset / == offset tests, equivalent for finite values).
template <int NP>
SW_DEV uint32_t mv_clip(const MV* in, uint32_t n, const float* plane, float offset, MV* out)
{
	auto dist_of = [&](const MV& v) {
		if constexpr(NP == 3)
			return (v.p[0] * plane[0] + v.p[1] * plane[1]) + v.p[2] * plane[2] - offset;
		else
			return v.p[0] * plane[0] + v.p[1] * plane[1] - offset;
	};
	uint32_t count = 0;
	const MV* prev = &in[n - 1];
	float prev_d = dist_of(*prev);
	for(uint32_t i = 0; i < n && count < MESH_MAX_POLY; ++i)
	{
		const MV& v = in[i];
		const float d = dist_of(v);
		if(d > 0.0f)
		{
			if(prev_d < 0.0f)
			{
				out[count++] = mv_intersect(*prev, v, prev_d, d, NP);
				if(count == MESH_MAX_POLY)
					break;
			}
			out[count++] = v;
		}
		else if(d == 0.0f)
			out[count++] = v;
		else if(prev_d > 0.0f)
			out[count++] = mv_intersect(*prev, v, prev_d, d, NP);
		prev = &v;
		prev_d = d;
	}
	return count;
}

// f32_to_fixed16 (fixed_math.rs): (x * 65536) as i32, saturating, NaN -> 0
SW_DEV int32_t mesh_fixed16(float x)
{
	const float y = x * 65536.0f;
	if(y != y)
		return 0;
	if(y >= 2147483647.0f)
		return INT32_MAX;
	if(y <= -2147483648.0f)
		return INT32_MIN;
	return int32_t(y);
}

struct ExpandLaunch
{
	const sw_mesh_run* runs;
	const uint64_t* bases; // per run: first vertex int, first fan-count byte
	const sw_model_vertex* vertices;
	const uint16_t* triangles;
	const uint32_t* counts;
	int32_t* out_ints;
	uint8_t* out_fans;
};

// Block (x, r): triangle slots [x * 256, x * 256 + 256) of run r, one per thread.
__global__ __launch_bounds__(256) void k_expand_mesh_runs(ExpandLaunch a)
{
	const sw_mesh_run& R = a.runs[blockIdx.y];
	const uint32_t t = blockIdx.x * 256u + threadIdx.x;
	if(t >= R.max_triangles)
		return;
	const uint64_t int_base = a.bases[blockIdx.y * 2], fan_base = a.bases[blockIdx.y * 2 + 1];
	const uint32_t visible = min(a.counts[R.count_index], R.max_triangles);
	uint32_t fan = 0;
	if(t < visible)
	{
		MV b0[MESH_MAX_POLY], b1[MESH_MAX_POLY];
		for(int k = 0; k < 3; ++k)
		{
			const uint32_t vi = min(uint32_t(a.triangles[(uint64_t(R.first_triangle) + t) * 3 + k]), R.num_vertices - 1u);
			const sw_model_vertex v = a.vertices[uint64_t(R.first_vertex) + vi];
			b0[k] = MV{{v.pos[0], v.pos[1], v.pos[2]}, {v.tc[0], v.tc[1]}, {v.light[0], v.light[1], v.light[2]}};
		}
		uint32_t n = 3;
		MV* src = b0;
		MV* dst = b1;
		for(uint32_t p = 0; p < R.num_planes3d && n >= 3; ++p)
		{
			n = mv_clip<3>(src, n, R.planes3d[p], R.planes3d[p][3], dst);
			MV* tmp = src;
			src = dst;
			dst = tmp;
		}
		if(n >= 3)
		{
			// pos.truncate() / pos.z
			for(uint32_t k = 0; k < n; ++k)
			{
				const float z = src[k].p[2];
				src[k].p[0] = src[k].p[0] / z;
				src[k].p[1] = src[k].p[1] / z;
				src[k].p[2] = 0.0f;
			}
			for(uint32_t p = 0; p < R.num_planes2d && n >= 3; ++p)
			{
				n = mv_clip<2>(src, n, R.planes2d[p], R.planes2d[p][2], dst);
				MV* tmp = src;
				src = dst;
				dst = tmp;
			}
		}
		if(n >= 3)
		{
			const uint32_t fan_max = 1u + R.num_planes3d + R.num_planes2d;
			fan = min(n - 2u, fan_max);
			int32_t* out = a.out_ints + int_base + uint64_t(t) * fan_max * 21u;
			for(uint32_t j = 0; j < fan; ++j)
			{
				const MV* tri[3] = {&src[0], &src[j + 1], &src[j + 2]};
				for(int v = 0; v < 3; ++v)
				{
					int32_t* o = out + j * 21u + v * 7;
					o[0] = mesh_fixed16(tri[v]->p[0]);
					o[1] = mesh_fixed16(tri[v]->p[1]);
					o[2] = mesh_fixed16(tri[v]->tc[0]);
					o[3] = mesh_fixed16(tri[v]->tc[1]);
					o[4] = mesh_fixed16(tri[v]->l[0]);
					o[5] = mesh_fixed16(tri[v]->l[1]);
					o[6] = mesh_fixed16(tri[v]->l[2]);
				}
			}
		}
	}
	a.out_fans[fan_base + t] = uint8_t(fan);
}

} // namespace

namespace swk
{
void launch_expand_mesh_runs(hipStream_t st, const sw_mesh_run* runs, uint32_t num_runs, uint32_t max_triangles, const uint64_t* bases,
	const sw_model_vertex* vertices, const uint16_t* triangles, const uint32_t* counts, int32_t* out_ints, uint8_t* out_fans)
{
	if(num_runs == 0 || max_triangles == 0)
		return;
	ExpandLaunch a{runs, bases, vertices, triangles, counts, out_ints, out_fans};
	hipLaunchKernelGGL(k_expand_mesh_runs, dim3((max_triangles + 255u) / 256u, num_runs), dim3(256), 0, st, a);
}
void launch_prepare_models(hipStream_t st, uint32_t num_instances, const sw_model_mesh* meshes, const sw_model_instance* instances,
	const uint64_t* sort_offsets, uint64_t* scratch, sw_model_vertex* out_vertices, uint16_t* out_triangles, uint32_t* out_counts,
	uint32_t* host_err)
{
	ModelLaunch L{meshes, instances, sort_offsets, scratch, out_vertices, out_triangles, out_counts, host_err};
	hipLaunchKernelGGL(k_prepare_models, dim3(num_instances), dim3(MODEL_BLOCK), 0, st, L);
}

uint32_t model_sort_lds_limit() { return MODEL_SORT_LDS; }
} // namespace swk
