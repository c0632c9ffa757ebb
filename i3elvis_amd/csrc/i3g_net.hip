/*
 * i3g warm-start network on MI355X: a 3D convolutional encoder-decoder
 * (levels of 16/32/64/64 channels, 3x3x3 kernels) that maps the node fields
 * (log10 viscosity, density anomaly) to a velocity-pressure guess for the
 * Stokes solve.  Every convolution -- forward, backward-data and
 * backward-weights -- is an implicit GEMM on the gfx950 matrix cores
 * (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation); weights are
 * kept in FP32 for Adam and packed to BF16 for the products.
 *
 * Layout: activations [D][H][W][C] over the node grid (z, x, y; the node
 * index of the Stokes fields is the pixel index), channels padded to a
 * multiple of 16.  Weights [27][Cout][Cin], tap t = 9 tz + 3 tx + ty.
 * C ABI in include/i3g.h.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/i3g.h"

#define NCHK(x)                                                                \
	do {                                                                   \
		hipError_t e_ = (x);                                           \
		if (e_ != hipSuccess) {                                        \
			fprintf(stderr, "i3g net: HIP error %s at %s:%d\n",   \
				hipGetErrorString(e_), __FILE__, __LINE__);    \
			return I3G_EDEVICE;                                    \
		}                                                              \
	} while (0)

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

#define NLAYER 7
#define WGRAD_CHUNK 1024 /* pixels per backward-weights partial */

static const int L_CIN[NLAYER] = {2, 16, 32, 64, 64, 32, 16};
static const int L_COUT[NLAYER] = {16, 32, 64, 64, 32, 16, 4};
static const int L_STRIDE[NLAYER] = {1, 2, 2, 2, 1, 1, 1};

static inline int pad16(int c) { return (c + 15) / 16 * 16; }

/* round to nearest even; inputs here are finite */
__device__ __forceinline__ bf16_t f2bf(float f)
{
	unsigned u = __float_as_uint(f);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(((unsigned)b) << 16); }

/* ------------------------------------------------------------------ */
/* Implicit-GEMM convolution (forward and backward-data)              */
/* ------------------------------------------------------------------ */

/* out[o, co] = bias[co] + sum_t sum_ci w[t][co][ci] * in(v(o, t), ci), with
 * v = o*stride + t - 1 per axis.  dil = 2 reads a zero-inserted input
 * (virtual coordinate v maps to v/2 when even, zero when odd): the
 * transposed convolution of a stride-2 layer.  One wavefront computes 16
 * output pixels x all output channels (CO16 tiles of 16).  A operand (16
 * pixels x 16 channels): lane l holds pixel l%16, channels 4(l/16)..+3; B
 * operand (16 channels x 16 outputs): lane l holds output channel l%16,
 * channels 4(l/16)..+3; result: lane l holds pixels 4(l/16)..+3 of output
 * channel l%16. */
template <int CO16>
__global__ void __launch_bounds__(256)
k_conv(int Do, int Ho, int Wo, int Di, int Hi, int Wi, int cinp, int coutp, int stride, int dil,
       const bf16_t *__restrict__ in, const bf16_t *__restrict__ w, const float *__restrict__ bias, int relu,
       int cout_store, bf16_t *out16, float *out32)
{
	int lane = threadIdx.x & 63;
	long pix0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
	long P = (long)Do * Ho * Wo;
	if (pix0 >= P) return;
	long p = pix0 + (lane & 15);
	int oz = 0, ox = 0, oy = 0;
	bool pv = p < P;
	if (pv) {
		oy = (int)(p % Wo);
		long r = p / Wo;
		ox = (int)(r % Ho);
		oz = (int)(r / Ho);
	}
	int kq = (lane >> 4) * 4;
	f4 acc[CO16];
	for (int j = 0; j < CO16; j++) acc[j] = (f4){0.f, 0.f, 0.f, 0.f};
	int nch = cinp / 16;
	for (int t = 0; t < 27; t++) {
		int tz = t / 9, tx = (t / 3) % 3, ty = t % 3;
		int vz = oz * stride + tz - 1, vx = ox * stride + tx - 1, vy = oy * stride + ty - 1;
		bool ok = pv;
		if (dil == 2) {
			ok = ok && vz >= 0 && vx >= 0 && vy >= 0 && !(vz & 1) && !(vx & 1) && !(vy & 1);
			vz >>= 1; vx >>= 1; vy >>= 1;
		}
		ok = ok && vz >= 0 && vx >= 0 && vy >= 0 && vz < Di && vx < Hi && vy < Wi;
		const bf16_t *src = in + (((long)vz * Hi + vx) * Wi + vy) * cinp + kq;
		const bf16_t *wt = w + (long)t * coutp * cinp + (long)(lane & 15) * cinp + kq;
		for (int c = 0; c < nch; c++) {
			s4 a = (s4){0, 0, 0, 0};
			if (ok) a = *(const s4 *)(src + c * 16);
			for (int j = 0; j < CO16; j++) {
				s4 b = *(const s4 *)(wt + (long)j * 16 * cinp + c * 16);
				acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
			}
		}
	}
	for (int j = 0; j < CO16; j++) {
		int co = j * 16 + (lane & 15);
		float bv = bias ? bias[co] : 0.f;
		for (int i = 0; i < 4; i++) {
			long q = pix0 + kq + i;
			if (q >= P) continue;
			float v = acc[j][i] + bv;
			if (relu) v = fmaxf(v, 0.f);
			if (out16) out16[q * coutp + co] = f2bf(v);
			if (out32 && co < cout_store) out32[q * cout_store + co] = v;
		}
	}
}

/* ------------------------------------------------------------------ */
/* Backward weights                                                   */
/* ------------------------------------------------------------------ */

/* Partial sums over one chunk of output pixels of
 *   dw[t][co][ci] = sum_o g[o][co] * in(v(o, t), ci)
 * as a 16 x 16 MFMA tile per wavefront (rows co, columns ci, reduction over
 * pixels, 16 per step: lane l holds pixels 4(l/16)..+3 of row/column l%16).
 * grid.x = chunk, grid.y = tap * (co tiles) * (ci tiles). */
__global__ void __launch_bounds__(64)
k_wgrad(int Do, int Ho, int Wo, int Di, int Hi, int Wi, int cinp, int coutp, int stride,
	const bf16_t *__restrict__ in, const bf16_t *__restrict__ g, float *part)
{
	int lane = threadIdx.x;
	int nct = coutp / 16, nit = cinp / 16;
	int y = blockIdx.y;
	int t = y / (nct * nit);
	int rem = y - t * nct * nit;
	int ct = rem / nit, it = rem - ct * nit;
	int tz = t / 9, tx = (t / 3) % 3, ty = t % 3;
	long P = (long)Do * Ho * Wo;
	long p0 = (long)blockIdx.x * WGRAD_CHUNK;
	long p1 = p0 + WGRAD_CHUNK < P ? p0 + WGRAD_CHUNK : P;
	int co = ct * 16 + (lane & 15), ci = it * 16 + (lane & 15);
	int kq = (lane >> 4) * 4;
	f4 acc = (f4){0.f, 0.f, 0.f, 0.f};
	for (long base = p0; base < p1; base += 16) {
		s4 a, b;
		for (int i = 0; i < 4; i++) {
			long p = base + kq + i;
			short av = 0, bv = 0;
			if (p < p1) {
				av = (short)g[p * coutp + co];
				int oy = (int)(p % Wo);
				long r = p / Wo;
				int ox = (int)(r % Ho), oz = (int)(r / Ho);
				int vz = oz * stride + tz - 1, vx = ox * stride + tx - 1, vy = oy * stride + ty - 1;
				if (vz >= 0 && vx >= 0 && vy >= 0 && vz < Di && vx < Hi && vy < Wi)
					bv = (short)in[(((long)vz * Hi + vx) * Wi + vy) * cinp + ci];
			}
			a[i] = av;
			b[i] = bv;
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
	}
	/* lane l: rows co_tile 4(l/16)+i, column ci_tile l%16 */
	float *dst = part + ((long)blockIdx.x * 27 + t) * coutp * cinp;
	for (int i = 0; i < 4; i++) {
		int r = ct * 16 + kq + i;
		dst[(long)r * cinp + ci] = acc[i];
	}
}

/* dw (fp32, logical [27][cout][cin]) += sum over chunks of the partials, in
 * chunk order; db[co] += sum_o g[o][co] from per-chunk sums */
__global__ void k_wgrad_reduce(long nchunk, int cinp, int coutp, int cin, int cout, const float *__restrict__ part,
			       float *dw)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long n = 27L * cout * cin;
	if (i >= n) return;
	int c = (int)(i % cin);
	long r = i / cin;
	int o = (int)(r % cout);
	int t = (int)(r / cout);
	long off = ((long)t * coutp + o) * cinp + c;
	long stride = 27L * coutp * cinp;
	float s = 0.f;
	for (long k = 0; k < nchunk; k++) s += part[k * stride + off];
	dw[i] += s;
}

__global__ void k_bgrad(long P, int coutp, int cout, const bf16_t *__restrict__ g, float *db)
{
	/* one block per output channel; fixed-order tree reduction */
	__shared__ float sh[256];
	int co = blockIdx.x;
	float s = 0.f;
	for (long p = threadIdx.x; p < P; p += 256) s += bf2f(g[p * coutp + co]);
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0 && co < cout) db[co] += sh[0];
}

/* ------------------------------------------------------------------ */
/* Elementwise pieces                                                 */
/* ------------------------------------------------------------------ */

/* node-linear upsampling (2n-1 <- n per axis) of a [D][H][W][C] bf16 field
 * plus a skip field: out = up(a) + skip */
__global__ void k_up_add(int D, int H, int W, int C, const bf16_t *__restrict__ a, const bf16_t *__restrict__ skip,
			 bf16_t *out)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long n = (long)D * H * W * C;
	if (i >= n) return;
	int c = (int)(i % C);
	long p = i / C;
	int y = (int)(p % W);
	long r = p / W;
	int x = (int)(r % H), z = (int)(r / H);
	int h = (H + 1) / 2, w = (W + 1) / 2;
	int z0 = z >> 1, z1 = (z + 1) >> 1, x0 = x >> 1, x1 = (x + 1) >> 1, y0 = y >> 1, y1 = (y + 1) >> 1;
	float s = 0.f;
	int zz[2] = {z0, z1}, xx[2] = {x0, x1}, yy[2] = {y0, y1};
	for (int q = 0; q < 8; q++) {
		int a0 = zz[q >> 2], a1 = xx[(q >> 1) & 1], a2 = yy[q & 1];
		s += bf2f(a[(((long)a0 * h + a1) * w + a2) * C + c]);
	}
	out[i] = f2bf(s * 0.125f + bf2f(skip[i]));
}

/* transpose of the upsampling: gout (fine, fp32) -> gin (coarse, fp32) */
__global__ void k_up_t(int D, int H, int W, int C, const float *__restrict__ gf, float *gc)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	int d = (D + 1) / 2, h = (H + 1) / 2, w = (W + 1) / 2;
	long n = (long)d * h * w * C;
	if (i >= n) return;
	int c = (int)(i % C);
	long p = i / C;
	int y = (int)(p % w);
	long r = p / w;
	int x = (int)(r % h), z = (int)(r / h);
	float s = 0.f;
	for (int dz = -1; dz <= 1; dz++) {
		int fz = 2 * z + dz;
		if (fz < 0 || fz >= D) continue;
		float wz = dz ? 0.5f : 1.f;
		for (int dx = -1; dx <= 1; dx++) {
			int fx = 2 * x + dx;
			if (fx < 0 || fx >= H) continue;
			float wx = dx ? 0.5f : 1.f;
			for (int dy = -1; dy <= 1; dy++) {
				int fy = 2 * y + dy;
				if (fy < 0 || fy >= W) continue;
				float wy = dy ? 0.5f : 1.f;
				s += wz * wx * wy * gf[(((long)fz * H + fx) * W + fy) * C + c];
			}
		}
	}
	gc[i] = s;
}

/* g = (g + add) * (act > 0) -> bf16 (add may be null); mask 0 keeps all */
__global__ void k_mask(long n, const float *__restrict__ g, const float *__restrict__ add,
		       const bf16_t *__restrict__ act, int mask, bf16_t *out, float *sum_out)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float v = g[i];
	if (add) v += add[i];
	if (sum_out) sum_out[i] = v;
	if (mask && !(bf2f(act[i]) > 0.f)) v = 0.f;
	out[i] = f2bf(v);
}

/* input [N][2] fp32 -> [N][16] bf16 */
__global__ void k_pack_in(long N, int cin, const float *__restrict__ x, bf16_t *out)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= N * 16) return;
	int c = (int)(i & 15);
	long p = i >> 4;
	out[i] = c < cin ? f2bf(x[p * cin + c]) : (bf16_t)0;
}

/* mean squared error against [N][4] targets: per-block partial loss and
 * the bf16 gradient 2 (y - t) / (4N) * scale into [N][16] */
__global__ void k_mse(long N, int cout, const float *__restrict__ y, const float *__restrict__ tg, float scale,
		      bf16_t *g, double *part)
{
	__shared__ double sh[256];
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double l = 0;
	if (i < N * 16) {
		int c = (int)(i & 15);
		long p = i >> 4;
		float gv = 0.f;
		if (c < cout) {
			float d = y[p * cout + c] - tg[p * cout + c];
			l = (double)d * d;
			gv = 2.f * d / (float)(N * cout) * scale;
		}
		g[i] = f2bf(gv);
	}
	sh[threadIdx.x] = l;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum(long n, const double *part, double *out)
{
	__shared__ double sh[256];
	double s = 0;
	for (long i = threadIdx.x; i < n; i += 256) s += part[i];
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) out[0] = sh[0];
}

/* Adam on the fp32 parameters (grad already averaged over the batch) */
__global__ void k_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float c1,
		       float c2, float eps)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	float gi = g[i];
	float mi = b1 * m[i] + (1.f - b1) * gi;
	float vi = b2 * v[i] + (1.f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* bf16 forward weights [27][coutp][cinp] and transposed, flipped weights
 * for the backward-data product [27][cinp][coutp] (tap 26-t) */
__global__ void k_pack_w(int cin, int cout, int cinp, int coutp, const float *__restrict__ w, bf16_t *wf, bf16_t *wb)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long n = 27L * coutp * cinp;
	if (i >= n) return;
	int c = (int)(i % cinp);
	long r = i / cinp;
	int o = (int)(r % coutp);
	int t = (int)(r / coutp);
	float v = (c < cin && o < cout) ? w[((long)t * cout + o) * cin + c] : 0.f;
	bf16_t b = f2bf(v);
	wf[i] = b;
	wb[((long)(26 - t) * cinp + c) * coutp + o] = b;
}

/* ------------------------------------------------------------------ */
/* Context                                                            */
/* ------------------------------------------------------------------ */

struct Layer {
	int cin, cout, cinp, coutp, stride;
	int Di, Hi, Wi, Do, Ho, Wo;   /* input / output grid */
	long woff, boff;              /* offsets in the fp32 parameter vector */
	bf16_t *wf = nullptr, *wb = nullptr;
};

struct i3g_net {
	hipStream_t s = nullptr;
	int n[3];                     /* level-0 grid (D=z, H=x, W=y) */
	int dims[4][3];
	Layer L[NLAYER];
	long nparam = 0;
	float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
	long step = 0;
	/* activations (bf16, channels padded) */
	bf16_t *x16 = nullptr, *e[4] = {}, *i3 = nullptr, *d3 = nullptr, *i2 = nullptr, *d2 = nullptr, *i1 = nullptr;
	float *y = nullptr, *tg = nullptr, *xin = nullptr;
	/* backward scratch */
	bf16_t *gz = nullptr;         /* bf16 operand of the current layer's backward */
	float *gA = nullptr, *gB = nullptr, *gC = nullptr;
	float *t0 = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
	float *part = nullptr;
	long partn = 0;
	double *lpart = nullptr, *lsum = nullptr;
};

template <typename T> static T *nalloc(size_t n, int *err)
{
	void *p = nullptr;
	if (n == 0) n = 1;
	if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { *err = I3G_ENOMEM; return nullptr; }
	return (T *)p;
}

static inline long vox(const int *d) { return (long)d[0] * d[1] * d[2]; }

extern "C" int i3g_warmstart_destroy(i3g_net *q)
{
	if (!q) return I3G_OK;
	if (q->s) hipStreamSynchronize(q->s);
	std::vector<void *> ps = {q->p, q->g, q->m, q->v, q->x16, q->i3, q->d3, q->i2, q->d2, q->i1, q->y, q->tg,
				  q->xin, q->gz, q->gA, q->gB, q->gC, q->t0, q->t1, q->t2, q->t3,
				  q->part, q->lpart, q->lsum};
	for (int k = 0; k < 4; k++) ps.push_back(q->e[k]);
	for (int l = 0; l < NLAYER; l++) { ps.push_back(q->L[l].wf); ps.push_back(q->L[l].wb); }
	for (void *p : ps)
		if (p) hipFree(p);
	if (q->s) hipStreamDestroy(q->s);
	delete q;
	return I3G_OK;
}

static int pack_weights(i3g_net *q)
{
	for (int l = 0; l < NLAYER; l++) {
		Layer &L = q->L[l];
		long n = 27L * L.coutp * L.cinp;
		hipLaunchKernelGGL(k_pack_w, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, q->s, L.cin, L.cout, L.cinp,
				   L.coutp, (const float *)(q->p + L.woff), L.wf, L.wb);
	}
	NCHK(hipGetLastError());
	return I3G_OK;
}

/* splitmix64 -> normal (Box-Muller) for He initialisation */
static double u01(unsigned long long *s)
{
	unsigned long long z = (*s += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	z ^= z >> 31;
	return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

extern "C" int i3g_warmstart_create(long nz, long nx, long ny, unsigned long long seed, i3g_net **out)
{
	if (!out) return I3G_EINVAL;
	long nn[3] = {nz, nx, ny};
	for (int a = 0; a < 3; a++)
		if (nn[a] < 9 || (nn[a] - 1) % 8 != 0 || nn[a] > 2049) return I3G_EINVAL;
	*out = nullptr;
	i3g_net *q = new i3g_net();
	for (int a = 0; a < 3; a++) {
		q->n[a] = (int)nn[a];
		q->dims[0][a] = (int)nn[a];
		for (int k = 1; k < 4; k++) q->dims[k][a] = (q->dims[k - 1][a] - 1) / 2 + 1;
	}
	int lev_in[NLAYER] = {0, 0, 1, 2, 2, 1, 0}, lev_out[NLAYER] = {0, 1, 2, 3, 2, 1, 0};
	long off = 0;
	for (int l = 0; l < NLAYER; l++) {
		Layer &L = q->L[l];
		L.cin = L_CIN[l]; L.cout = L_COUT[l]; L.stride = L_STRIDE[l];
		L.cinp = pad16(L.cin); L.coutp = pad16(L.cout);
		L.Di = q->dims[lev_in[l]][0]; L.Hi = q->dims[lev_in[l]][1]; L.Wi = q->dims[lev_in[l]][2];
		L.Do = q->dims[lev_out[l]][0]; L.Ho = q->dims[lev_out[l]][1]; L.Wo = q->dims[lev_out[l]][2];
		L.woff = off; off += 27L * L.cin * L.cout;
		L.boff = off; off += L.cout;
	}
	q->nparam = off;
	int err = 0;
	if (hipStreamCreate(&q->s) != hipSuccess) { delete q; return I3G_EDEVICE; }
	q->p = nalloc<float>(off, &err); q->g = nalloc<float>(off, &err);
	q->m = nalloc<float>(off, &err); q->v = nalloc<float>(off, &err);
	for (int l = 0; l < NLAYER; l++) {
		q->L[l].wf = nalloc<bf16_t>(27L * q->L[l].coutp * q->L[l].cinp, &err);
		q->L[l].wb = nalloc<bf16_t>(27L * q->L[l].coutp * q->L[l].cinp, &err);
	}
	long P0 = vox(q->dims[0]), P1 = vox(q->dims[1]), P2 = vox(q->dims[2]), P3 = vox(q->dims[3]);
	q->x16 = nalloc<bf16_t>(P0 * 16, &err);
	q->e[0] = nalloc<bf16_t>(P0 * 16, &err); q->e[1] = nalloc<bf16_t>(P1 * 32, &err);
	q->e[2] = nalloc<bf16_t>(P2 * 64, &err); q->e[3] = nalloc<bf16_t>(P3 * 64, &err);
	q->i3 = nalloc<bf16_t>(P2 * 64, &err); q->d3 = nalloc<bf16_t>(P2 * 32, &err);
	q->i2 = nalloc<bf16_t>(P1 * 32, &err); q->d2 = nalloc<bf16_t>(P1 * 16, &err);
	q->i1 = nalloc<bf16_t>(P0 * 16, &err);
	q->y = nalloc<float>(P0 * 4, &err); q->tg = nalloc<float>(P0 * 4, &err); q->xin = nalloc<float>(P0 * 2, &err);
	long big = std::max(P0 * 16, std::max(P1 * 32, P2 * 64));
	q->gz = nalloc<bf16_t>(big, &err);
	q->gA = nalloc<float>(P0 * 16, &err); q->gB = nalloc<float>(P1 * 32, &err); q->gC = nalloc<float>(P2 * 64, &err);
	q->t0 = nalloc<float>(P0 * 16, &err); q->t1 = nalloc<float>(P1 * 32, &err); q->t2 = nalloc<float>(P2 * 64, &err);
	q->t3 = nalloc<float>(P3 * 64, &err);
	long pn = 0;
	for (int l = 0; l < NLAYER; l++) {
		Layer &L = q->L[l];
		long ch = ((long)L.Do * L.Ho * L.Wo + WGRAD_CHUNK - 1) / WGRAD_CHUNK;
		pn = std::max(pn, ch * 27L * L.coutp * L.cinp);
	}
	q->part = nalloc<float>(pn, &err);
	q->partn = pn;
	q->lpart = nalloc<double>((P0 * 16 + 255) / 256, &err);
	q->lsum = nalloc<double>(1, &err);
	if (err) { i3g_warmstart_destroy(q); return err; }
	/* He-normal weights, zero biases */
	std::vector<float> hp(off, 0.f);
	unsigned long long st = seed * 0x2545F4914F6CDD1DULL + 1;
	for (int l = 0; l < NLAYER; l++) {
		Layer &L = q->L[l];
		double sd = std::sqrt(2.0 / (27.0 * L.cin));
		for (long i = 0; i < 27L * L.cin * L.cout; i++) {
			double a = u01(&st), b = u01(&st);
			hp[L.woff + i] = (float)(sd * std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b));
		}
	}
	if (hipMemcpy(q->p, hp.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(q->m, 0, off * sizeof(float)) != hipSuccess || hipMemset(q->v, 0, off * sizeof(float)) != hipSuccess) {
		i3g_warmstart_destroy(q);
		return I3G_EDEVICE;
	}
	int rc = pack_weights(q);
	if (rc) { i3g_warmstart_destroy(q); return rc; }
	*out = q;
	return I3G_OK;
}

extern "C" long i3g_warmstart_nparams(const i3g_net *q) { return q ? q->nparam : I3G_EINVAL; }

extern "C" int i3g_warmstart_get_params(i3g_net *q, float *p)
{
	if (!q || !p) return I3G_EINVAL;
	NCHK(hipMemcpyAsync(p, q->p, q->nparam * sizeof(float), hipMemcpyDeviceToHost, q->s));
	NCHK(hipStreamSynchronize(q->s));
	return I3G_OK;
}

extern "C" int i3g_warmstart_set_params(i3g_net *q, const float *p)
{
	if (!q || !p) return I3G_EINVAL;
	NCHK(hipMemcpyAsync(q->p, p, q->nparam * sizeof(float), hipMemcpyHostToDevice, q->s));
	int rc = pack_weights(q);
	if (rc) return rc;
	NCHK(hipStreamSynchronize(q->s));
	return I3G_OK;
}

static dim3 lin(long n) { return dim3((unsigned)((n + 255) / 256)); }

static void conv_launch(i3g_net *q, const Layer &L, int Do, int Ho, int Wo, int Di, int Hi, int Wi, int cinp,
			int coutp, int stride, int dil, const bf16_t *in, const bf16_t *w, const float *bias, int relu,
			int cout_store, bf16_t *o16, float *o32)
{
	(void)L;
	long P = (long)Do * Ho * Wo;
	dim3 grid((unsigned)((P + 63) / 64));
	switch (coutp / 16) {
	case 1:
		hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, q->s, Do, Ho, Wo, Di, Hi, Wi, cinp, coutp, stride, dil, in, w,
				   bias, relu, cout_store, o16, o32);
		break;
	case 2:
		hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, q->s, Do, Ho, Wo, Di, Hi, Wi, cinp, coutp, stride, dil, in, w,
				   bias, relu, cout_store, o16, o32);
		break;
	default:
		hipLaunchKernelGGL(k_conv<4>, grid, dim3(256), 0, q->s, Do, Ho, Wo, Di, Hi, Wi, cinp, coutp, stride, dil, in, w,
				   bias, relu, cout_store, o16, o32);
		break;
	}
}

/* forward pass from q->xin (fp32 [P0][2]) to q->y (fp32 [P0][4]) */
static int forward(i3g_net *q)
{
	long P0 = vox(q->dims[0]);
	hipLaunchKernelGGL(k_pack_in, lin(P0 * 16), dim3(256), 0, q->s, P0, 2, (const float *)q->xin, q->x16);
	const bf16_t *src = q->x16;
	for (int l = 0; l < 4; l++) {
		Layer &L = q->L[l];
		conv_launch(q, L, L.Do, L.Ho, L.Wo, L.Di, L.Hi, L.Wi, L.cinp, L.coutp, L.stride, 1, src, L.wf, q->p + L.boff,
			    1, 0, q->e[l], nullptr);
		src = q->e[l];
	}
	/* decoder: i3 = up(e4) + e3, d3 = relu(conv), i2 = up(d3) + e2, ... */
	const int *d2 = q->dims[2], *d1 = q->dims[1], *d0 = q->dims[0];
	hipLaunchKernelGGL(k_up_add, lin(vox(d2) * 64), dim3(256), 0, q->s, d2[0], d2[1], d2[2], 64, (const bf16_t *)q->e[3],
			   (const bf16_t *)q->e[2], q->i3);
	Layer &L4 = q->L[4];
	conv_launch(q, L4, L4.Do, L4.Ho, L4.Wo, L4.Di, L4.Hi, L4.Wi, L4.cinp, L4.coutp, 1, 1, q->i3, L4.wf, q->p + L4.boff, 1,
		    0, q->d3, nullptr);
	hipLaunchKernelGGL(k_up_add, lin(vox(d1) * 32), dim3(256), 0, q->s, d1[0], d1[1], d1[2], 32, (const bf16_t *)q->d3,
			   (const bf16_t *)q->e[1], q->i2);
	Layer &L5 = q->L[5];
	conv_launch(q, L5, L5.Do, L5.Ho, L5.Wo, L5.Di, L5.Hi, L5.Wi, L5.cinp, L5.coutp, 1, 1, q->i2, L5.wf, q->p + L5.boff, 1,
		    0, q->d2, nullptr);
	hipLaunchKernelGGL(k_up_add, lin(vox(d0) * 16), dim3(256), 0, q->s, d0[0], d0[1], d0[2], 16, (const bf16_t *)q->d2,
			   (const bf16_t *)q->e[0], q->i1);
	Layer &L6 = q->L[6];
	conv_launch(q, L6, L6.Do, L6.Ho, L6.Wo, L6.Di, L6.Hi, L6.Wi, L6.cinp, L6.coutp, 1, 1, q->i1, L6.wf, q->p + L6.boff, 0,
		    4, nullptr, q->y);
	NCHK(hipGetLastError());
	return I3G_OK;
}

/* gradient of conv layer l given the bf16 operand q->gz (dL/d pre-activation
 * at the layer output): accumulates dW, db into q->g and, when gin is not
 * null, writes dL/d input (fp32) */
static int layer_backward(i3g_net *q, int l, const bf16_t *in, float *gin)
{
	Layer &L = q->L[l];
	long P = (long)L.Do * L.Ho * L.Wo;
	long nch = (P + WGRAD_CHUNK - 1) / WGRAD_CHUNK;
	dim3 gw((unsigned)nch, (unsigned)(27 * (L.coutp / 16) * (L.cinp / 16)));
	hipLaunchKernelGGL(k_wgrad, gw, dim3(64), 0, q->s, L.Do, L.Ho, L.Wo, L.Di, L.Hi, L.Wi, L.cinp, L.coutp, L.stride,
			   in, (const bf16_t *)q->gz, q->part);
	long nw = 27L * L.cout * L.cin;
	hipLaunchKernelGGL(k_wgrad_reduce, lin(nw), dim3(256), 0, q->s, nch, L.cinp, L.coutp, L.cin, L.cout,
			   (const float *)q->part, q->g + L.woff);
	hipLaunchKernelGGL(k_bgrad, dim3(L.coutp), dim3(256), 0, q->s, P, L.coutp, L.cout, (const bf16_t *)q->gz,
			   q->g + L.boff);
	if (gin) {
		/* transposed convolution: stride 1 over the (zero-inserted when the
		 * layer strides) output gradient with flipped, transposed weights */
		conv_launch(q, L, L.Di, L.Hi, L.Wi, L.Do, L.Ho, L.Wo, L.coutp, L.cinp, 1, L.stride, (const bf16_t *)q->gz,
			    L.wb, nullptr, 0, L.cinp, nullptr, gin);
	}
	NCHK(hipGetLastError());
	return I3G_OK;
}

/* backward pass after forward(); q->tg holds the target; adds the
 * gradient (times scale) into q->g; the summed squared error goes to
 * q->lsum.  Buffers: gA [P0][16] (d i1 = d e1 from the skip), gB [P1][32]
 * (d i2), gC [P2][64] (d i3), tN the coarse-level scratch of level N. */
static int backward(i3g_net *q, float scale)
{
	const int *dm0 = q->dims[0], *dm1 = q->dims[1], *dm2 = q->dims[2], *dm3 = q->dims[3];
	long P0 = vox(dm0), P1 = vox(dm1), P2 = vox(dm2), P3 = vox(dm3);
	long nb = (P0 * 16 + 255) / 256;
	hipLaunchKernelGGL(k_mse, dim3((unsigned)nb), dim3(256), 0, q->s, P0, 4, (const float *)q->y, (const float *)q->tg,
			   scale, q->gz, q->lpart);
	hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, q->s, nb, (const double *)q->lpart, q->lsum);
	int rc;
	/* decoder */
	if ((rc = layer_backward(q, 6, q->i1, q->gA))) return rc;
	hipLaunchKernelGGL(k_up_t, lin(P1 * 16), dim3(256), 0, q->s, dm0[0], dm0[1], dm0[2], 16, (const float *)q->gA, q->t1);
	hipLaunchKernelGGL(k_mask, lin(P1 * 16), dim3(256), 0, q->s, P1 * 16, (const float *)q->t1, (const float *)nullptr,
			   (const bf16_t *)q->d2, 1, q->gz, (float *)nullptr);
	if ((rc = layer_backward(q, 5, q->i2, q->gB))) return rc;
	hipLaunchKernelGGL(k_up_t, lin(P2 * 32), dim3(256), 0, q->s, dm1[0], dm1[1], dm1[2], 32, (const float *)q->gB, q->t2);
	hipLaunchKernelGGL(k_mask, lin(P2 * 32), dim3(256), 0, q->s, P2 * 32, (const float *)q->t2, (const float *)nullptr,
			   (const bf16_t *)q->d3, 1, q->gz, (float *)nullptr);
	if ((rc = layer_backward(q, 4, q->i3, q->gC))) return rc;
	/* encoder: d e4 = up_t(d i3) */
	hipLaunchKernelGGL(k_up_t, lin(P3 * 64), dim3(256), 0, q->s, dm2[0], dm2[1], dm2[2], 64, (const float *)q->gC, q->t3);
	hipLaunchKernelGGL(k_mask, lin(P3 * 64), dim3(256), 0, q->s, P3 * 64, (const float *)q->t3, (const float *)nullptr,
			   (const bf16_t *)q->e[3], 1, q->gz, (float *)nullptr);
	if ((rc = layer_backward(q, 3, q->e[2], q->t2))) return rc;
	/* d e3 = conv-back + skip part (gC) */
	hipLaunchKernelGGL(k_mask, lin(P2 * 64), dim3(256), 0, q->s, P2 * 64, (const float *)q->t2, (const float *)q->gC,
			   (const bf16_t *)q->e[2], 1, q->gz, (float *)nullptr);
	if ((rc = layer_backward(q, 2, q->e[1], q->t1))) return rc;
	hipLaunchKernelGGL(k_mask, lin(P1 * 32), dim3(256), 0, q->s, P1 * 32, (const float *)q->t1, (const float *)q->gB,
			   (const bf16_t *)q->e[1], 1, q->gz, (float *)nullptr);
	if ((rc = layer_backward(q, 1, q->e[0], q->t0))) return rc;
	hipLaunchKernelGGL(k_mask, lin(P0 * 16), dim3(256), 0, q->s, P0 * 16, (const float *)q->t0, (const float *)q->gA,
			   (const bf16_t *)q->e[0], 1, q->gz, (float *)nullptr);
	if ((rc = layer_backward(q, 0, q->x16, nullptr))) return rc;
	NCHK(hipGetLastError());
	return I3G_OK;
}

extern "C" int i3g_warmstart_infer(i3g_net *q, const float *in, float *out)
{
	if (!q || !in || !out) return I3G_EINVAL;
	long P0 = vox(q->dims[0]);
	NCHK(hipMemcpyAsync(q->xin, in, P0 * 2 * sizeof(float), hipMemcpyHostToDevice, q->s));
	int rc = forward(q);
	if (rc) return rc;
	NCHK(hipMemcpyAsync(out, q->y, P0 * 4 * sizeof(float), hipMemcpyDeviceToHost, q->s));
	NCHK(hipStreamSynchronize(q->s));
	return I3G_OK;
}

extern "C" int i3g_warmstart_infer_device(i3g_net *q, long n, const float *d_in, float *d_out)
{
	if (!q || !d_in || !d_out) return I3G_EINVAL;
	long P0 = vox(q->dims[0]);
	if (n != P0) return I3G_EINVAL;
	NCHK(hipMemcpyAsync(q->xin, d_in, P0 * 2 * sizeof(float), hipMemcpyDeviceToDevice, q->s));
	int rc = forward(q);
	if (rc) return rc;
	NCHK(hipMemcpyAsync(d_out, q->y, P0 * 4 * sizeof(float), hipMemcpyDeviceToDevice, q->s));
	NCHK(hipStreamSynchronize(q->s));
	return I3G_OK;
}

extern "C" int i3g_warmstart_grad(i3g_net *q, const float *in, const float *target, double *loss, float *grad)
{
	if (!q || !in || !target) return I3G_EINVAL;
	long P0 = vox(q->dims[0]);
	NCHK(hipMemsetAsync(q->g, 0, q->nparam * sizeof(float), q->s));
	NCHK(hipMemcpyAsync(q->xin, in, P0 * 2 * sizeof(float), hipMemcpyHostToDevice, q->s));
	NCHK(hipMemcpyAsync(q->tg, target, P0 * 4 * sizeof(float), hipMemcpyHostToDevice, q->s));
	int rc = forward(q);
	if (!rc) rc = backward(q, 1.f);
	if (rc) return rc;
	double ls = 0;
	NCHK(hipMemcpyAsync(&ls, q->lsum, sizeof(double), hipMemcpyDeviceToHost, q->s));
	if (grad) NCHK(hipMemcpyAsync(grad, q->g, q->nparam * sizeof(float), hipMemcpyDeviceToHost, q->s));
	NCHK(hipStreamSynchronize(q->s));
	if (loss) *loss = ls / (double)(P0 * 4);
	return I3G_OK;
}

extern "C" int i3g_warmstart_train_step(i3g_net *q, int batch, const float *in, const float *target, double lr,
					double *loss)
{
	if (!q || !in || !target || batch < 1 || !(lr > 0)) return I3G_EINVAL;
	long P0 = vox(q->dims[0]);
	NCHK(hipMemsetAsync(q->g, 0, q->nparam * sizeof(float), q->s));
	double tot = 0;
	for (int b = 0; b < batch; b++) {
		NCHK(hipMemcpyAsync(q->xin, in + (long)b * P0 * 2, P0 * 2 * sizeof(float), hipMemcpyHostToDevice, q->s));
		NCHK(hipMemcpyAsync(q->tg, target + (long)b * P0 * 4, P0 * 4 * sizeof(float), hipMemcpyHostToDevice, q->s));
		int rc = forward(q);
		if (!rc) rc = backward(q, 1.f / (float)batch);
		if (rc) return rc;
		double ls = 0;
		NCHK(hipMemcpyAsync(&ls, q->lsum, sizeof(double), hipMemcpyDeviceToHost, q->s));
		NCHK(hipStreamSynchronize(q->s));
		tot += ls / (double)(P0 * 4);
	}
	q->step++;
	const float b1 = 0.9f, b2 = 0.999f;
	float c1 = 1.f - powf(b1, (float)q->step), c2 = 1.f - powf(b2, (float)q->step);
	hipLaunchKernelGGL(k_adam, lin(q->nparam), dim3(256), 0, q->s, q->nparam, q->p, (const float *)q->g, q->m, q->v,
			   (float)lr, b1, b2, c1, c2, 1e-8f);
	int rc = pack_weights(q);
	if (rc) return rc;
	NCHK(hipStreamSynchronize(q->s));
	if (loss) *loss = tot / batch;
	return I3G_OK;
}
