// af_image.h -- DM-time images for the _dedisperse_spec return type.
//
// Follows the reference's preprocessing (include/data.h:619-727): the plane
// is cut into slices of t_sample seconds; each slice (DM rows x samples, as
// float32) is resized to 512 x 512 -- area averaging when both sides shrink,
// an 8-tap Lanczos (a = 4) kernel otherwise -- min-max normalised to 0..255
// and mapped through the viridis colour map (BGR bytes).  The reference does
// this with OpenCV (cv::resize INTER_AREA / INTER_LANCZOS4, cv::normalize
// NORM_MINMAX, cv::applyColorMap COLORMAP_VIRIDIS); this is an independent
// implementation of those definitions.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "viridis_lut.h"

namespace afimg {

// For each destination index, the source indices and weights that form it.
struct Taps {
  std::vector<int> start;   // [dsize + 1] offsets into idx / w
  std::vector<int> idx;
  std::vector<float> w;
};

// Area averaging for ssize > dsize: destination cell dx covers source
// [dx*s, (dx+1)*s), s = ssize/dsize; each source pixel is weighted by its
// overlap with the cell, divided by the cell width.
inline Taps area_taps(int ssize, int dsize) {
  Taps t;
  const double s = (double)ssize / dsize;
  t.start.push_back(0);
  for (int dx = 0; dx < dsize; ++dx) {
    const double f1 = dx * s, f2 = f1 + s;
    const double cell = std::min(s, ssize - f1);
    int s1 = (int)std::ceil(f1), s2 = (int)std::floor(f2);
    s2 = std::min(s2, ssize - 1);
    s1 = std::min(s1, s2);
    if (s1 - f1 > 1e-3) {
      t.idx.push_back(s1 - 1);
      t.w.push_back((float)((s1 - f1) / cell));
    }
    for (int sx = s1; sx < s2; ++sx) {
      t.idx.push_back(sx);
      t.w.push_back((float)(1.0 / cell));
    }
    if (f2 - s2 > 1e-3) {
      t.idx.push_back(s2);
      t.w.push_back((float)(std::min(std::min(f2 - s2, 1.0), cell) / cell));
    }
    t.start.push_back((int)t.idx.size());
  }
  return t;
}

// Lanczos a = 4: destination centre dx maps to source x = (dx + 0.5)*s - 0.5;
// 8 taps floor(x)-3 .. floor(x)+4, weights sinc(u) sinc(u/4) normalised to
// sum 1; taps outside the source repeat the edge sample.
inline Taps lanczos4_taps(int ssize, int dsize) {
  Taps t;
  const double s = (double)ssize / dsize;
  const double pi = 3.14159265358979323846;
  t.start.push_back(0);
  for (int dx = 0; dx < dsize; ++dx) {
    const double fx = (dx + 0.5) * s - 0.5;
    const int sx = (int)std::floor(fx);
    const double fr = fx - sx;
    double w[8], sum = 0;
    for (int k = 0; k < 8; ++k) {
      const double u = (k - 3) - fr;
      if (std::fabs(u) < 1e-12) {
        w[k] = 1.0;
      } else {
        w[k] = std::sin(pi * u) * std::sin(pi * u / 4.0) / (pi * pi * u * u / 4.0);
      }
      sum += w[k];
    }
    for (int k = 0; k < 8; ++k) {
      t.idx.push_back(std::min(std::max(sx - 3 + k, 0), ssize - 1));
      t.w.push_back((float)(w[k] / sum));
    }
    t.start.push_back((int)t.idx.size());
  }
  return t;
}

// One slice -> 512 x 512 x 3 BGR bytes.  src(r, c) = plane[r * ld + c0 + c]
// for c < cols.
template <typename V>
void slice_image(const V *plane, size_t ld, size_t c0, int rows, int cols, int size,
                 uint8_t *out) {
  const bool area = rows > size && cols > size;
  const Taps tx = area ? area_taps(cols, size) : lanczos4_taps(cols, size);
  const Taps ty = area ? area_taps(rows, size) : lanczos4_taps(rows, size);
  std::vector<float> tmp((size_t)rows * size), img((size_t)size * size);
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r) {
    const V *src = plane + (size_t)r * ld + c0;
    float *o = tmp.data() + (size_t)r * size;
    for (int dx = 0; dx < size; ++dx) {
      float acc = 0.0f;
      for (int k = tx.start[dx]; k < tx.start[dx + 1]; ++k) acc += tx.w[k] * (float)src[tx.idx[k]];
      o[dx] = acc;
    }
  }
#pragma omp parallel for schedule(static)
  for (int dy = 0; dy < size; ++dy) {
    float *o = img.data() + (size_t)dy * size;
    for (int dx = 0; dx < size; ++dx) o[dx] = 0.0f;
    for (int k = ty.start[dy]; k < ty.start[dy + 1]; ++k) {
      const float wy = ty.w[k];
      const float *row = tmp.data() + (size_t)ty.idx[k] * size;
      for (int dx = 0; dx < size; ++dx) o[dx] += wy * row[dx];
    }
  }
  float mn = img[0], mx = img[0];
  for (float v : img) {
    mn = std::min(mn, v);
    mx = std::max(mx, v);
  }
  // NORM_MINMAX to [0, 255]: scale 0 when the slice is flat
  const double scale = (mx - mn) > 2.220446049250313e-16 ? 255.0 / ((double)mx - mn) : 0.0;
  const float fs = (float)scale, fo = (float)(-(double)mn * scale);
  for (size_t i = 0; i < img.size(); ++i) {
    float v = std::nearbyint(img[i] * fs + fo);   // round half to even
    v = std::min(std::max(v, 0.0f), 255.0f);
    const uint8_t *c = kViridisBGR[(int)v];
    out[3 * i + 0] = c[0];
    out[3 * i + 1] = c[1];
    out[3 * i + 2] = c[2];
  }
}

}  // namespace afimg
