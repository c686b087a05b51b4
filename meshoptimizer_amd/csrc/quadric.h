// Device-side quadric algebra of the edge-collapse simplifier.
//
// Each function keeps the expression tree of the reference routine it names
// (src/simplifier.cpp), term for term and in the same association order, so that with
// -ffp-contract=off every result rounds exactly as the reference's scalar x86 code does.
#pragma once

#include "common.h"

namespace mhip
{

struct V3
{
	float x, y, z;
};

__device__ __forceinline__ V3 load_pos(const float4* pos, uint32_t i)
{
	float4 p = pos[i];
	return V3{p.x, p.y, p.z};
}

__device__ __forceinline__ Quad load_quad(const Quad* q, uint32_t i)
{
	const float4* p = reinterpret_cast<const float4*>(q + i);
	float4 a = p[0], b = p[1], c = p[2];
	Quad r;
	r.a00 = a.x, r.a11 = a.y, r.a22 = a.z, r.a10 = a.w;
	r.a20 = b.x, r.a21 = b.y, r.b0 = b.z, r.b1 = b.w;
	r.b2 = c.x, r.c = c.y, r.w = c.z, r.pad = 0.f;
	return r;
}

__device__ __forceinline__ void store_quad(Quad* q, uint32_t i, const Quad& r)
{
	float4* p = reinterpret_cast<float4*>(q + i);
	p[0] = make_float4(r.a00, r.a11, r.a22, r.a10);
	p[1] = make_float4(r.a20, r.a21, r.b0, r.b1);
	p[2] = make_float4(r.b2, r.c, r.w, 0.f);
}

__device__ __forceinline__ void quad_zero(Quad& Q)
{
	Q.a00 = Q.a11 = Q.a22 = Q.a10 = 0.f;
	Q.a20 = Q.a21 = Q.b0 = Q.b1 = 0.f;
	Q.b2 = Q.c = Q.w = Q.pad = 0.f;
}

// quadricAdd, src/simplifier.cpp:716-729
__device__ __forceinline__ void quad_add(Quad& Q, const Quad& R)
{
	Q.a00 += R.a00;
	Q.a11 += R.a11;
	Q.a22 += R.a22;
	Q.a10 += R.a10;
	Q.a20 += R.a20;
	Q.a21 += R.a21;
	Q.b0 += R.b0;
	Q.b1 += R.b1;
	Q.b2 += R.b2;
	Q.c += R.c;
	Q.w += R.w;
}

__device__ __forceinline__ void grad_add(Grad& G, const Grad& R)
{
	G.gx += R.gx;
	G.gy += R.gy;
	G.gz += R.gz;
	G.gw += R.gw;
}

// normalize, src/simplifier.cpp:702-714
__device__ __forceinline__ float v3_normalize(V3& v)
{
	float length = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);

	if (length > 0)
	{
		v.x /= length;
		v.y /= length;
		v.z /= length;
	}

	return length;
}

// quadricEval, src/simplifier.cpp:750-774
__device__ __forceinline__ float quad_eval(const Quad& Q, const V3& v)
{
	float rx = Q.b0;
	float ry = Q.b1;
	float rz = Q.b2;

	rx += Q.a10 * v.y;
	ry += Q.a21 * v.z;
	rz += Q.a20 * v.x;

	rx *= 2;
	ry *= 2;
	rz *= 2;

	rx += Q.a00 * v.x;
	ry += Q.a11 * v.y;
	rz += Q.a22 * v.z;

	float r = Q.c;
	r += rx * v.x;
	r += ry * v.y;
	r += rz * v.z;
	return r;
}

// quadricError (position), src/simplifier.cpp:776-782
__device__ __forceinline__ float quad_error(const Quad& Q, const V3& v)
{
	float r = quad_eval(Q, v);
	float s = Q.w == 0.f ? 0.f : 1.f / Q.w;
	return fabsf(r) * s;
}

// quadricError (attributes), src/simplifier.cpp:784-799
__device__ __forceinline__ float quad_error_attr(const Quad& Q, const Grad* G, int n, const V3& v, const float* va)
{
	float r = quad_eval(Q, v);

	for (int k = 0; k < n; ++k)
	{
		float a = va[k];
		Grad g4 = G[k];
		float g = v.x * g4.gx + v.y * g4.gy + v.z * g4.gz + g4.gw;
		r += a * (a * Q.w - 2 * g);
	}

	return fabsf(r);
}

// quadricFromPlane, src/simplifier.cpp:801-819
__device__ __forceinline__ void quad_from_plane(Quad& Q, float a, float b, float c, float d, float w)
{
	float aw = a * w;
	float bw = b * w;
	float cw = c * w;
	float dw = d * w;

	Q.a00 = a * aw;
	Q.a11 = b * bw;
	Q.a22 = c * cw;
	Q.a10 = a * bw;
	Q.a20 = a * cw;
	Q.a21 = b * cw;
	Q.b0 = a * dw;
	Q.b1 = b * dw;
	Q.b2 = c * dw;
	Q.c = d * dw;
	Q.w = w;
	Q.pad = 0.f;
}

// quadricFromPoint, src/simplifier.cpp:821-830
__device__ __forceinline__ void quad_from_point(Quad& Q, float x, float y, float z, float w)
{
	Q.a00 = Q.a11 = Q.a22 = w;
	Q.a10 = Q.a20 = Q.a21 = 0;
	Q.b0 = -x * w;
	Q.b1 = -y * w;
	Q.b2 = -z * w;
	Q.c = (x * x + y * y + z * z) * w;
	Q.w = w;
	Q.pad = 0.f;
}

// quadricFromTriangle, src/simplifier.cpp:832-845
__device__ __forceinline__ void quad_from_triangle(Quad& Q, const V3& p0, const V3& p1, const V3& p2, float weight)
{
	V3 p10 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
	V3 p20 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};

	V3 normal = {p10.y * p20.z - p10.z * p20.y, p10.z * p20.x - p10.x * p20.z, p10.x * p20.y - p10.y * p20.x};
	float area = v3_normalize(normal);

	float distance = normal.x * p0.x + normal.y * p0.y + normal.z * p0.z;

	quad_from_plane(Q, normal.x, normal.y, normal.z, -distance, sqrtf(area) * weight);
}

// quadricFromTriangleEdge, src/simplifier.cpp:847-868
__device__ __forceinline__ void quad_from_triangle_edge(Quad& Q, const V3& p0, const V3& p1, const V3& p2, float weight)
{
	V3 p10 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};

	float lengthsq = p10.x * p10.x + p10.y * p10.y + p10.z * p10.z;
	float length = sqrtf(lengthsq);

	V3 p20 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
	float p20p = p20.x * p10.x + p20.y * p10.y + p20.z * p10.z;

	V3 perp = {p20.x * lengthsq - p10.x * p20p, p20.y * lengthsq - p10.y * p20p, p20.z * lengthsq - p10.z * p20p};
	v3_normalize(perp);

	float distance = perp.x * p0.x + perp.y * p0.y + perp.z * p0.z;

	quad_from_plane(Q, perp.x, perp.y, perp.z, -distance, length * weight);
}

// the edge quadric that fillEdgeQuadrics adds, src/simplifier.cpp:1141-1153
__device__ __forceinline__ void quad_edge_term(Quad& Q, const V3& p0, const V3& p1, const V3& p2, float edge_weight)
{
	quad_from_triangle_edge(Q, p0, p1, p2, edge_weight);

	Quad QT;
	quad_from_triangle(QT, p0, p1, p2, edge_weight);
	QT.w = 0;
	quad_add(Q, QT);
}

// quadricFromAttributes shared factors, src/simplifier.cpp:870-909
struct AttrFrame
{
	float w;
	float gx1, gx2, gy1, gy2, gz1, gz2;
	V3 p0;
};

__device__ __forceinline__ AttrFrame attr_frame(const V3& p0, const V3& p1, const V3& p2)
{
	V3 p10 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
	V3 p20 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};

	V3 normal = {p10.y * p20.z - p10.z * p20.y, p10.z * p20.x - p10.x * p20.z, p10.x * p20.y - p10.y * p20.x};
	float area = sqrtf(normal.x * normal.x + normal.y * normal.y + normal.z * normal.z) * 0.5f;

	const V3& v0 = p10;
	const V3& v1 = p20;
	float d00 = v0.x * v0.x + v0.y * v0.y + v0.z * v0.z;
	float d01 = v0.x * v1.x + v0.y * v1.y + v0.z * v1.z;
	float d11 = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
	float denom = d00 * d11 - d01 * d01;
	float denomr = denom == 0 ? 0.f : 1.f / denom;

	AttrFrame f;
	f.w = area;
	f.gx1 = (d11 * v0.x - d01 * v1.x) * denomr;
	f.gx2 = (d00 * v1.x - d01 * v0.x) * denomr;
	f.gy1 = (d11 * v0.y - d01 * v1.y) * denomr;
	f.gy2 = (d00 * v1.y - d01 * v0.y) * denomr;
	f.gz1 = (d11 * v0.z - d01 * v1.z) * denomr;
	f.gz2 = (d00 * v1.z - d01 * v0.z) * denomr;
	f.p0 = p0;
	return f;
}

// one attribute's gradient over the triangle before the area weight, src/simplifier.cpp:915-926
__device__ __forceinline__ Grad attr_gradient(const AttrFrame& f, float a0, float a1, float a2)
{
	float gx = f.gx1 * (a1 - a0) + f.gx2 * (a2 - a0);
	float gy = f.gy1 * (a1 - a0) + f.gy2 * (a2 - a0);
	float gz = f.gz1 * (a1 - a0) + f.gz2 * (a2 - a0);
	float gw = a0 - f.p0.x * gx - f.p0.y * gy - f.p0.z * gz;
	return Grad{gx, gy, gz, gw};
}

// the weighted gradient alone: what attr_term gives as g
__device__ __forceinline__ Grad attr_weighted_gradient(const AttrFrame& f, float a0, float a1, float a2)
{
	const Grad r = attr_gradient(f, a0, a1, a2);
	const float w = f.w;
	return Grad{w * r.gx, w * r.gy, w * r.gz, w * r.gw};
}

// one attribute's gradient and its quadric terms, src/simplifier.cpp:915-948.
// Q accumulates the attribute terms in k order; g receives the weighted gradient.
__device__ __forceinline__ void attr_term(Quad& Q, Grad& g, const AttrFrame& f, float a0, float a1, float a2)
{
	const Grad r = attr_gradient(f, a0, a1, a2);
	const float gx = r.gx, gy = r.gy, gz = r.gz, gw = r.gw;

	float w = f.w;
	Q.a00 += w * (gx * gx);
	Q.a11 += w * (gy * gy);
	Q.a22 += w * (gz * gz);

	Q.a10 += w * (gy * gx);
	Q.a20 += w * (gz * gx);
	Q.a21 += w * (gz * gy);

	Q.b0 += w * (gx * gw);
	Q.b1 += w * (gy * gw);
	Q.b2 += w * (gz * gw);

	Q.c += w * (gw * gw);

	g.gx = w * gx;
	g.gy = w * gy;
	g.gz = w * gz;
	g.gw = w * gw;
}

// quadricVolumeGradient, src/simplifier.cpp:952-965: unit normal scaled by half the area, and
// the plane offset through p0 with the same scale
__device__ __forceinline__ void grad_volume(Grad& G, const V3& p0, const V3& p1, const V3& p2)
{
	V3 e1 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
	V3 e2 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
	V3 n = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
	float half_area = v3_normalize(n) * 0.5f;

	G.gx = n.x * half_area;
	G.gy = n.y * half_area;
	G.gz = n.z * half_area;
	G.gw = (-p0.x * n.x - p0.y * n.y - p0.z * n.z) * half_area;
}

// hasTriangleFlip, src/simplifier.cpp:1182-1198
__device__ __forceinline__ bool triangle_flips(const V3& a, const V3& b, const V3& c, const V3& d)
{
	V3 eb = {b.x - a.x, b.y - a.y, b.z - a.z};
	V3 ec = {c.x - a.x, c.y - a.y, c.z - a.z};
	V3 ed = {d.x - a.x, d.y - a.y, d.z - a.z};

	V3 nbc = {eb.y * ec.z - eb.z * ec.y, eb.z * ec.x - eb.x * ec.z, eb.x * ec.y - eb.y * ec.x};
	V3 nbd = {eb.y * ed.z - eb.z * ed.y, eb.z * ed.x - eb.x * ed.z, eb.x * ed.y - eb.y * ed.x};

	float ndp = nbc.x * nbd.x + nbc.y * nbd.y + nbc.z * nbd.z;
	float abc = nbc.x * nbc.x + nbc.y * nbc.y + nbc.z * nbc.z;
	float abd = nbd.x * nbd.x + nbd.y * nbd.y + nbd.z * nbd.z;

	return ndp <= 0.25f * sqrtf(abc * abd);
}

// sortEdgeCollapses key, src/simplifier.cpp:1456-1468: 8 exponent + 4 mantissa bits, clamped
__device__ __forceinline__ uint32_t collapse_sort_key(float error)
{
	uint32_t key = (__float_as_uint(error) << 1) >> 20;
	return key < 2560u ? key : 2559u;
}

} // namespace mhip
