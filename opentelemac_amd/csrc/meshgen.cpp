// Seeded variable-radius Poisson-disk point sampling (Bridson's algorithm)
// for the synthetic unstructured test meshes. Host-only helper of the test
// and benchmark harness: it generates inputs and is not on the solve path.
//
// The sampling domain is given as a polygon (any simple polygon); the local
// minimum spacing is r0 * factor(x, y), where factor is interpolated from a
// regular grid the caller supplies (so the caller controls refinement zones).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {  // splitmix64: small, seedable, platform independent
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

bool inside(const double* px, const double* py, int n, double x, double y) {
  bool c = false;
  for (int i = 0, j = n - 1; i < n; j = i++) {
    if (((py[i] > y) != (py[j] > y)) &&
        (x < (px[j] - px[i]) * (y - py[i]) / (py[j] - py[i]) + px[i]))
      c = !c;
  }
  return c;
}

}  // namespace

extern "C" {

// Returns the number of points written (at most max_points) or -1 when the
// buffer was too small. fx is a (gny, gnx) row-major grid of spacing factors
// over the polygon's bounding box.
// n_init fixed points (e.g. samples of the outline) are inserted first,
// without distance test; sampling then grows from them.
long meshgen_poisson_disk(const double* px, const double* py, int npoly,
                          double r0, const double* fgrid, int gnx, int gny,
                          uint64_t seed, int k_attempts, const double* init_x,
                          const double* init_y, long n_init, double* out_x,
                          double* out_y, long max_points) {
  double xmin = px[0], xmax = px[0], ymin = py[0], ymax = py[0];
  for (int i = 1; i < npoly; ++i) {
    xmin = std::fmin(xmin, px[i]);
    xmax = std::fmax(xmax, px[i]);
    ymin = std::fmin(ymin, py[i]);
    ymax = std::fmax(ymax, py[i]);
  }
  double fmin_ = 1e300, fmax_ = 0;
  for (long i = 0; i < (long)gnx * gny; ++i) {
    fmin_ = std::fmin(fmin_, fgrid[i]);
    fmax_ = std::fmax(fmax_, fgrid[i]);
  }
  const double rmin = r0 * fmin_, rmax = r0 * fmax_;
  auto radius = [&](double x, double y) {
    double gx = (x - xmin) / (xmax - xmin) * (gnx - 1);
    double gy = (y - ymin) / (ymax - ymin) * (gny - 1);
    int ix = (int)std::floor(gx), iy = (int)std::floor(gy);
    ix = ix < 0 ? 0 : (ix > gnx - 2 ? gnx - 2 : ix);
    iy = iy < 0 ? 0 : (iy > gny - 2 ? gny - 2 : iy);
    double tx = gx - ix, ty = gy - iy;
    tx = std::fmin(std::fmax(tx, 0.0), 1.0);
    ty = std::fmin(std::fmax(ty, 0.0), 1.0);
    double f = (1 - tx) * (1 - ty) * fgrid[(long)iy * gnx + ix] +
               tx * (1 - ty) * fgrid[(long)iy * gnx + ix + 1] +
               (1 - tx) * ty * fgrid[(long)(iy + 1) * gnx + ix] +
               tx * ty * fgrid[(long)(iy + 1) * gnx + ix + 1];
    return r0 * f;
  };
  // coarse inside/outside raster. Cells an edge passes through, and their
  // eight neighbours, are tested exactly (2); every other cell lies wholly
  // on one side of the outline and takes the status of its centre, found by
  // a scan of each row's edge crossings.
  const int mnx = 1024, mny = 1024;
  const double mdx = (xmax - xmin) / mnx, mdy = (ymax - ymin) / mny;
  std::vector<signed char> mask((size_t)mnx * mny, 0);
  {
    auto mark = [&](double x, double y) {
      const int i = (int)std::floor((x - xmin) / mdx), j = (int)std::floor((y - ymin) / mdy);
      for (int jj = j - 1; jj <= j + 1; ++jj)
        for (int ii = i - 1; ii <= i + 1; ++ii)
          if (ii >= 0 && jj >= 0 && ii < mnx && jj < mny) mask[(size_t)jj * mnx + ii] = 2;
    };
    const double step = 0.25 * std::fmin(mdx, mdy);
    for (int a = 0, b = npoly - 1; a < npoly; b = a++) {
      const double ex = px[a] - px[b], ey = py[a] - py[b];
      const long ns = (long)std::ceil(std::sqrt(ex * ex + ey * ey) / step) + 1;
      for (long k = 0; k <= ns; ++k) {
        const double t = (double)k / (double)ns;
        mark(px[b] + t * ex, py[b] + t * ey);
      }
    }
    std::vector<double> xs;
    for (int j = 0; j < mny; ++j) {
      const double cy = ymin + (j + 0.5) * mdy;
      xs.clear();
      for (int a = 0, b = npoly - 1; a < npoly; b = a++)
        if ((py[a] > cy) != (py[b] > cy))
          xs.push_back((px[b] - px[a]) * (cy - py[a]) / (py[b] - py[a]) + px[a]);
      for (int i = 0; i < mnx; ++i) {
        signed char& v = mask[(size_t)j * mnx + i];
        if (v == 2) continue;
        const double cx = xmin + (i + 0.5) * mdx;
        // even-odd: the number of crossings right of the centre
        int cnt = 0;
        for (double xc : xs) cnt += cx < xc;
        v = (cnt & 1) ? 1 : 0;
      }
    }
  }
  auto in_domain = [&](double x, double y) {
    int i = (int)((x - xmin) / mdx), j = (int)((y - ymin) / mdy);
    if (i < 0 || j < 0 || i >= mnx || j >= mny) return false;
    signed char v = mask[(size_t)j * mnx + i];
    if (v != 2) return v == 1;
    return inside(px, py, npoly, x, y);
  };
  const double cell = rmin / std::sqrt(2.0);
  const long ncx = (long)std::ceil((xmax - xmin) / cell) + 1;
  const long ncy = (long)std::ceil((ymax - ymin) / cell) + 1;
  std::vector<std::vector<int>> grid((size_t)(ncx * ncy));
  const int reach = (int)std::ceil(rmax / cell);
  Rng rng(seed);
  long n = 0;
  std::vector<long> active;
  std::vector<double> rad_of;
  rad_of.reserve((size_t)max_points);
  auto add = [&](double x, double y) -> bool {
    if (n >= max_points) return false;
    out_x[n] = x;
    out_y[n] = y;
    rad_of.push_back(radius(x, y));
    long cx = (long)((x - xmin) / cell), cy = (long)((y - ymin) / cell);
    cx = cx < 0 ? 0 : (cx >= ncx ? ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= ncy ? ncy - 1 : cy);
    grid[(size_t)(cy * ncx + cx)].push_back((int)n);
    active.push_back(n);
    ++n;
    return true;
  };
  auto ok = [&](double x, double y, double r) {
    long cx = (long)((x - xmin) / cell), cy = (long)((y - ymin) / cell);
    for (long j = cy - reach; j <= cy + reach; ++j) {
      if (j < 0 || j >= ncy) continue;
      for (long i = cx - reach; i <= cx + reach; ++i) {
        if (i < 0 || i >= ncx) continue;
        for (int q : grid[(size_t)(j * ncx + i)]) {
          double dx = out_x[q] - x, dy = out_y[q] - y;
          double rq = std::fmax(r, rad_of[q]);
          if (dx * dx + dy * dy < rq * rq) return false;
        }
      }
    }
    return true;
  };
  for (long i = 0; i < n_init; ++i)
    if (!add(init_x[i], init_y[i])) return -1;
  // seed point: first random point inside the polygon
  for (int tries = 0; tries < 100000 && n_init == 0; ++tries) {
    double x = xmin + rng.uniform() * (xmax - xmin);
    double y = ymin + rng.uniform() * (ymax - ymin);
    if (in_domain(x, y)) {
      add(x, y);
      break;
    }
  }
  // active list processed in FIFO order (grows as a front: locality of ids)
  size_t head = 0;
  while (head < active.size()) {
    long p = active[head];
    double x0 = out_x[p], y0 = out_y[p];
    double r = rad_of[p];
    bool found = false;
    for (int k = 0; k < k_attempts; ++k) {
      double ang = 2.0 * M_PI * rng.uniform();
      double rad = r * (1.0 + rng.uniform());
      double x = x0 + rad * std::cos(ang), y = y0 + rad * std::sin(ang);
      if (x < xmin || x > xmax || y < ymin || y > ymax) continue;
      if (!in_domain(x, y)) continue;
      double rx = radius(x, y);
      if (ok(x, y, rx)) {
        if (!add(x, y)) return -1;
        found = true;
      }
    }
    if (!found) ++head;
  }
  return n;
}

}  // extern "C"
