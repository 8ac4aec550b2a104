// af_image.hip -- DM-time images of a device plane.  Compiled with
// -ffp-contract=off.
//
// Replaces the host-side preprocessing of include/data.h:619-727: the plane
// is cut into slices of `sps` columns; each slice (DM rows x samples, as
// float32) is resized to size x size -- area averaging when both sides
// shrink, an 8-tap Lanczos (a = 4) kernel otherwise -- min-max normalised to
// 0..255 and mapped through the viridis colour map (BGR bytes).
//
// The resize weights are af_image.h's, built on the host and uploaded.  Each
// kernel adds a pixel's taps in the order af_image.h does, with the same
// float operations (no contraction), so the images equal the host module's
// bit for bit.  Four launches per batch of slices:
//   k_img_rows   tmp[s][r][dx] = sum_k wx[k] * (float)plane[r][s*sps + ix[k]]
//   k_img_cols   img[s][dy][dx] = sum_k wy[k] * tmp[s][iy[k]][dx]
//   k_img_minmax per slice (one workgroup per slice)
//   k_img_map    rint(img * scale + offset), clamped, through the colour table
// The plane is read once per slice (its columns are disjoint), the images
// are written once; everything else (about 2 KB of tap tables, the
// rows x size and size x size float intermediates) is small.
#include "af_internal.h"
#include "af_image.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace af {
namespace {

struct TapsDev {
  const int *start;   // [size + 1]
  const int *idx;
  const float *w;
};

template <typename V>
__global__ __launch_bounds__(256) void k_img_rows(const V *__restrict__ plane, size_t ld,
                                                  size_t sps, int rows, int size, int s0,
                                                  int last, TapsDev tx, TapsDev txl,
                                                  float *__restrict__ tmp) {
  const int s = s0 + (int)blockIdx.y;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)rows * (size_t)size) return;
  const int r = (int)(t / (size_t)size), dx = (int)(t % (size_t)size);
  const TapsDev T = s == last ? txl : tx;
  const V *src = plane + (size_t)r * ld + (size_t)s * sps;
  float acc = 0.0f;
  for (int k = T.start[dx]; k < T.start[dx + 1]; ++k) acc += T.w[k] * (float)src[T.idx[k]];
  tmp[((size_t)blockIdx.y * rows + r) * size + dx] = acc;
}

__global__ __launch_bounds__(256) void k_img_cols(const float *__restrict__ tmp, int rows,
                                                  int size, int s0, int last, TapsDev ty,
                                                  TapsDev tyl, float *__restrict__ img) {
  const int s = s0 + (int)blockIdx.y;
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)size * (size_t)size) return;
  const int dy = (int)(t / (size_t)size), dx = (int)(t % (size_t)size);
  const TapsDev T = s == last ? tyl : ty;
  const float *base = tmp + (size_t)blockIdx.y * rows * size + dx;
  float acc = 0.0f;
  for (int k = T.start[dy]; k < T.start[dy + 1]; ++k)
    acc += T.w[k] * base[(size_t)T.idx[k] * size];
  img[((size_t)blockIdx.y * size + dy) * size + dx] = acc;
}

// one workgroup per slice: its smallest and largest value
__global__ __launch_bounds__(1024) void k_img_minmax(const float *__restrict__ img, int npix,
                                                     float2 *__restrict__ mm) {
  __shared__ float smn[1024], smx[1024];
  const float *p = img + (size_t)blockIdx.x * npix;
  float mn = p[0], mx = p[0];
  for (int i = threadIdx.x; i < npix; i += 1024) {
    mn = fminf(mn, p[i]);
    mx = fmaxf(mx, p[i]);
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + h]);
      smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) mm[blockIdx.x] = make_float2(smn[0], smx[0]);
}

// NORM_MINMAX to [0, 255] (scale 0 for a flat slice), round half to even,
// clamp, colour table (BGR)
__global__ __launch_bounds__(256) void k_img_map(const float *__restrict__ img, int npix,
                                                 const float2 *__restrict__ mm,
                                                 const uint8_t *__restrict__ lut,
                                                 uint8_t *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)npix) return;
  const float2 m = mm[blockIdx.y];
  const double scale = (m.y - m.x) > 2.220446049250313e-16 ? 255.0 / ((double)m.y - m.x) : 0.0;
  const float fs = (float)scale, fo = (float)(-(double)m.x * scale);
  const size_t j = (size_t)blockIdx.y * npix + i;
  float v = rintf(img[j] * fs + fo);
  v = fminf(fmaxf(v, 0.0f), 255.0f);
  const uint8_t *c = lut + 3 * (int)v;
  out[3 * j + 0] = c[0];
  out[3 * j + 1] = c[1];
  out[3 * j + 2] = c[2];
}

// float intermediates of one batch of slices stay below this
constexpr size_t kImgBatchBytes = size_t(256) << 20;

template <typename V>
hipError_t launch_rows(const void *plane, size_t ld, size_t sps, int rows, int size, int s0,
                       int nb, int last, TapsDev tx, TapsDev txl, float *tmp, hipStream_t st) {
  const unsigned gx = (unsigned)(((size_t)rows * size + 255) / 256);
  hipLaunchKernelGGL(k_img_rows<V>, dim3(gx, (unsigned)nb), dim3(256), 0, st,
                     (const V *)plane, ld, sps, rows, size, s0, last, tx, txl, tmp);
  return hipGetLastError();
}

}  // namespace

af_status image_slices(const void *d_plane, int dtype, size_t rows, size_t cols, size_t ld,
                       size_t sps, int size, uint8_t *d_images, hipStream_t st, void **keep,
                       size_t *keep_n) {
  if (!d_plane || !d_images) return fail(AF_ERR_INVALID, "null argument");
  if (dtype != AF_U32 && dtype != AF_U64 && dtype != AF_F32)
    return fail(AF_ERR_INVALID, "plane element type must be AF_U32, AF_U64 or AF_F32");
  if (rows == 0 || cols == 0 || sps == 0 || size <= 0 || size > 8192 || ld < cols)
    return fail(AF_ERR_INVALID, "empty plane, zero slice width, bad image size or ld < cols");
  if (rows > (size_t)INT32_MAX / (size_t)size || rows * (size_t)size > ((size_t)1 << 36))
    return fail(AF_ERR_INVALID, "plane too tall for the image kernels");
  const size_t nsl = (cols + sps - 1) / sps;
  const size_t wfull = std::min(sps, cols);
  const size_t wlast = cols - (nsl - 1) * sps;
  if (wfull > (size_t)INT32_MAX) return fail(AF_ERR_INVALID, "slice too wide");
  const int R = (int)rows;
  // the resize taps of a full slice and of the last one (af_image.h)
  auto build = [&](int w, afimg::Taps &tx, afimg::Taps &ty) {
    const bool area = R > size && w > size;
    tx = area ? afimg::area_taps(w, size) : afimg::lanczos4_taps(w, size);
    ty = area ? afimg::area_taps(R, size) : afimg::lanczos4_taps(R, size);
  };
  afimg::Taps t[4];
  build((int)wfull, t[0], t[1]);
  build((int)wlast, t[2], t[3]);
  // one upload: per table start, idx, w; then the colour table
  std::vector<uint8_t> blob;
  size_t off[4][3];
  auto put = [&](const void *p, size_t n) {
    const size_t o = (blob.size() + 15) & ~(size_t)15;
    blob.resize(o + n);
    if (n) std::memcpy(blob.data() + o, p, n);
    return o;
  };
  for (int k = 0; k < 4; ++k) {
    off[k][0] = put(t[k].start.data(), t[k].start.size() * sizeof(int));
    off[k][1] = put(t[k].idx.data(), t[k].idx.size() * sizeof(int));
    off[k][2] = put(t[k].w.data(), t[k].w.size() * sizeof(float));
  }
  const size_t off_lut = put(&kViridisBGR[0][0], 256 * 3);
  const size_t per = ((size_t)R * size + (size_t)size * size) * sizeof(float) + sizeof(float2);
  const size_t nb = std::max<size_t>(1, std::min<size_t>({nsl, 65535, kImgBatchBytes / per}));
  const size_t tmp_b = nb * (size_t)R * size * sizeof(float);
  const size_t img_b = nb * (size_t)size * size * sizeof(float);
  const size_t blob_o = (tmp_b + img_b + nb * sizeof(float2) + 255) & ~(size_t)255;
  // scratch: the caller's kept buffer (grown when too small) or one of its own
  uint8_t *scratch = nullptr;
  const size_t need = blob_o + blob.size();
  if (keep && *keep && *keep_n >= need) {
    scratch = (uint8_t *)*keep;
  } else {
    if (keep && *keep) {
      (void)hipFree(*keep);
      *keep = nullptr;
      *keep_n = 0;
    }
    if (hipMalloc(&scratch, need) != hipSuccess) {
      (void)hipGetLastError();
      return fail(AF_ERR_NOMEM, "image scratch allocation failed");
    }
    if (keep) {
      *keep = scratch;
      *keep_n = need;
    }
  }
  float *tmp = (float *)scratch;
  float *img = (float *)(scratch + tmp_b);
  float2 *mm = (float2 *)(scratch + tmp_b + img_b);
  uint8_t *dblob = scratch + blob_o;
  hipError_t e = hipMemcpyAsync(dblob, blob.data(), blob.size(), hipMemcpyHostToDevice, st);
  auto taps = [&](int k) {
    return TapsDev{(const int *)(dblob + off[k][0]), (const int *)(dblob + off[k][1]),
                   (const float *)(dblob + off[k][2])};
  };
  const TapsDev tx = taps(0), ty = taps(1), txl = taps(2), tyl = taps(3);
  const int last = (int)(nsl - 1);
  const int npix = size * size;
  for (size_t s0 = 0; s0 < nsl && e == hipSuccess; s0 += nb) {
    const int n = (int)std::min(nb, nsl - s0);
    switch (dtype) {
    case AF_U32:
      e = launch_rows<uint32_t>(d_plane, ld, sps, R, size, (int)s0, n, last, tx, txl, tmp, st);
      break;
    case AF_U64:
      e = launch_rows<uint64_t>(d_plane, ld, sps, R, size, (int)s0, n, last, tx, txl, tmp, st);
      break;
    default:
      e = launch_rows<float>(d_plane, ld, sps, R, size, (int)s0, n, last, tx, txl, tmp, st);
      break;
    }
    if (e != hipSuccess) break;
    const unsigned gp = (unsigned)(((size_t)npix + 255) / 256);
    hipLaunchKernelGGL(k_img_cols, dim3(gp, (unsigned)n), dim3(256), 0, st, tmp, R, size,
                       (int)s0, last, ty, tyl, img);
    hipLaunchKernelGGL(k_img_minmax, dim3((unsigned)n), dim3(1024), 0, st, img, npix, mm);
    hipLaunchKernelGGL(k_img_map, dim3(gp, (unsigned)n), dim3(256), 0, st, img, npix, mm,
                       dblob + off_lut, d_images + s0 * (size_t)npix * 3);
    e = hipGetLastError();
  }
  // the host tables and the scratch live until the work is done
  const hipError_t es = hipStreamSynchronize(st);
  if (e == hipSuccess) e = es;
  if (!keep) (void)hipFree(scratch);
  if (e != hipSuccess) return fail(AF_ERR_RUNTIME, std::string("imaging: ") + hipGetErrorString(e));
  return AF_OK;
}

}  // namespace af

extern "C" af_status af_image_slices(const void *d_plane, int plane_dtype, size_t rows,
                                     size_t cols, size_t ld, size_t sps, int size,
                                     uint8_t *d_images, void *stream) {
  return af::image_slices(d_plane, plane_dtype, rows, cols, ld, sps, size, d_images,
                          (hipStream_t)stream, nullptr, nullptr);
}
