// ugrid_cut.cpp -- cut-cell geometry and SRD neighbourhoods of a closed solid polygon on the
// tensor-product grid (ugrid_cut_build in ugrid_hip.h): the steps of the reference's
// cut2d_build_polygon (src/cutcell2d.cpp:247-503) on this project's polygon primitives
// (cutcell.h), which follow the reference's helpers operation for operation, so every array is
// bit-identical to the reference's:
//   1. every cell (ghosts included) full fluid: area dx dy, centroid the cell centre;
//   2. x / y face apertures (fluid length and centroid along the face), faces outside the
//      polygon's bounding box fully open, snapped to the full width / zero within 1e-13;
//   3. per cell the clipped wall segments (fluid on the left), the area and first moments from
//      boundary integrals over the walls and the open faces, the wall vector and centroid; a
//      cell without walls is solid when its centre lies inside the polygon; an area below
//      1e-13 dx dy is solid;
//   4. faces next to a solid cell closed;
//   5. the body at least three cells from the domain boundary;
//   6. SRD neighbourhoods of the cells below half their full area: merge with the neighbour
//      along the wall normal when that reaches half the area, else the 3 x 3, else the 5 x 5
//      block of fluid cells; neighbourhood counts, weighted volumes and centroids.
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "cutcell.h"
#include "ugrid_hip.h"

struct ugrid_cut {
  int nx = 0, ny = 0, mx = 0, my = 0;
  int nsmall = 0, nbad = 0, nsolid = 0, ncut = 0;
  std::vector<double> vol, cx, cy, flag, axl, axy, ayl, ayx, wax, way, wcx, wcy, nbrcount, smallpos;
  std::vector<double> sidx, scnt, smem, svhat, sxhat, syhat;
  std::vector<amr_wall> walls;
};

namespace {

int report(char* err, size_t len, int code, const std::string& message) {
  if (err && len) {
    const size_t k = message.size() < len - 1 ? message.size() : len - 1;
    std::memcpy(err, message.data(), k);
    err[k] = 0;
  }
  return code;
}

int build(const ugrid_axes& a, const double* px, const double* py, int n, const int32_t* patches, ugrid_cut& c,
          char* err, size_t err_len) {
  using cfd::hipamr::Polygon;
  if (n < 3) return report(err, err_len, 1, "cut geometry: the polygon needs three vertices");
  for (int q = 0; q < n; ++q) {
    const int e = (q + 1) % n;
    if (!std::isfinite(px[q]) || !std::isfinite(py[q]) || (px[q] == px[e] && py[q] == py[e]) ||
        (patches && patches[q] < 0))
      return report(err, err_len, 1, "cut geometry: invalid polygon vertex or patch");
  }
  const int nx = a.nx, ny = a.ny, ng = 2, mx = nx + 4, my = ny + 4;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int q = 0; q < n; ++q) {
    const int e = (q + 1) % n;
    if (px[q] == px[e])
      for (int i = 0; i <= nx; ++i)
        if (std::fabs(px[q] - a.xf[i]) <= 16 * eps * std::fmax(1.0, std::fabs(px[q])))
          return report(err, err_len, 3, "cut geometry: edge " + std::to_string(q) + " lies on an x grid line");
    if (py[q] == py[e])
      for (int j = 0; j <= ny; ++j)
        if (std::fabs(py[q] - a.yf[j]) <= 16 * eps * std::fmax(1.0, std::fabs(py[q])))
          return report(err, err_len, 3, "cut geometry: edge " + std::to_string(q) + " lies on a y grid line");
  }
  const Polygon body{px, py, n, patches};
  c.nx = nx;
  c.ny = ny;
  c.mx = mx;
  c.my = my;
  const size_t cells = static_cast<size_t>(mx) * my;
  const size_t nfx = static_cast<size_t>(nx + 1) * ny, nfy = static_cast<size_t>(nx) * (ny + 1);
  for (auto* v : {&c.vol, &c.cx, &c.cy, &c.flag, &c.wax, &c.way, &c.wcx, &c.wcy}) v->assign(cells, 0.0);
  c.nbrcount.assign(cells, 1.0);
  c.smallpos.assign(cells, -1.0);
  for (auto* v : {&c.axl, &c.axy}) v->assign(nfx, 0.0);
  for (auto* v : {&c.ayl, &c.ayx}) v->assign(nfy, 0.0);

  // 1. all cells full fluid
  for (int j = -ng; j < ny + ng; ++j)
    for (int i = -ng; i < nx + ng; ++i) {
      const size_t k = static_cast<size_t>(j + ng) * mx + (i + ng);
      c.vol[k] = a.dx[i + ng] * a.dy[j + ng];
      c.cx[k] = a.xc[i + ng];
      c.cy[k] = a.yc[j + ng];
    }
  double hmin = 1e300;
  for (int i = 0; i < nx; ++i) hmin = std::fmin(hmin, a.dx[i + ng]);
  for (int j = 0; j < ny; ++j) hmin = std::fmin(hmin, a.dy[j + ng]);
  double x_lo = px[0], x_hi = px[0], y_lo = py[0], y_hi = py[0];
  for (int q = 1; q < n; ++q) {
    x_lo = std::fmin(x_lo, px[q]);
    x_hi = std::fmax(x_hi, px[q]);
    y_lo = std::fmin(y_lo, py[q]);
    y_hi = std::fmax(y_hi, py[q]);
  }

  // 2. apertures
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const size_t f = static_cast<size_t>(j) * (nx + 1) + i;
      const double h = a.dy[j + ng];
      if (a.xf[i] < x_lo || a.xf[i] > x_hi || a.yf[j + 1] < y_lo || a.yf[j] > y_hi) {
        c.axl[f] = h;
        c.axy[f] = a.yc[j + ng];
        continue;
      }
      double len, cen;
      cfd::hipamr::open_aperture_raw(body, 0, a.xf[i], a.yf[j], a.yf[j + 1], &len, &cen);
      c.axl[f] = len;
      c.axy[f] = cen;
      if (std::fabs(len - h) < 1e-13 * h) c.axl[f] = h, c.axy[f] = a.yc[j + ng];
      if (len < 1e-13 * h) c.axl[f] = 0.0, c.axy[f] = a.yc[j + ng];
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t f = static_cast<size_t>(j) * nx + i;
      const double h = a.dx[i + ng];
      if (a.yf[j] < y_lo || a.yf[j] > y_hi || a.xf[i + 1] < x_lo || a.xf[i] > x_hi) {
        c.ayl[f] = h;
        c.ayx[f] = a.xc[i + ng];
        continue;
      }
      double len, cen;
      cfd::hipamr::open_aperture_raw(body, 1, a.yf[j], a.xf[i], a.xf[i + 1], &len, &cen);
      c.ayl[f] = len;
      c.ayx[f] = cen;
      if (std::fabs(len - h) < 1e-13 * h) c.ayl[f] = h, c.ayx[f] = a.xc[i + ng];
      if (len < 1e-13 * h) c.ayl[f] = 0.0, c.ayx[f] = a.xc[i + ng];
    }

  // 3. walls, areas and centroids of the cells the polygon's box touches
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t k = static_cast<size_t>(j + ng) * mx + (i + ng);
      const double x0 = a.xf[i], x1 = a.xf[i + 1], y0 = a.yf[j], y1 = a.yf[j + 1];
      if (x1 < x_lo || x0 > x_hi || y1 < y_lo || y0 > y_hi) continue;
      double wvx = 0.0, wvy = 0.0, wlen = 0.0, wmx = 0.0, wmy = 0.0;
      double A = 0.0, MX = 0.0, MY = 0.0;  // A = int x dy, MX = 1/2 int x^2 dy, MY = -1/2 int y^2 dx
      bool walled = false;
      const size_t first_wall = c.walls.size();
      for (int e = 0, b0 = n - 1; e < n; b0 = e++) {  // edge b0 -> e
        double sx1 = px[b0], sy1 = py[b0], sx2 = px[e], sy2 = py[e];
        if (std::fmax(sx1, sx2) < x0 || std::fmin(sx1, sx2) > x1 || std::fmax(sy1, sy2) < y0 ||
            std::fmin(sy1, sy2) > y1)
          continue;
        if (!cfd::hipamr::clip_to_rect(&sx1, &sy1, &sx2, &sy2, x0, x1, y0, y1)) continue;
        double nfx, nfy;
        cfd::hipamr::fluid_normal(body, b0, 0.5 * (px[b0] + px[e]), 0.5 * (py[b0] + py[e]), 1e-7 * hmin, &nfx, &nfy);
        const double len = std::sqrt((sx2 - sx1) * (sx2 - sx1) + (sy2 - sy1) * (sy2 - sy1));
        const double tx = nfy, ty = -nfx;  // fluid on the left
        if ((sx2 - sx1) * tx + (sy2 - sy1) * ty < 0.0) {
          std::swap(sx1, sx2);
          std::swap(sy1, sy2);
        }
        A += 0.5 * (sx1 + sx2) * (sy2 - sy1);
        MX += (sx1 * sx1 + sx1 * sx2 + sx2 * sx2) * (sy2 - sy1) / 6.0;
        MY += -(sy1 * sy1 + sy1 * sy2 + sy2 * sy2) * (sx2 - sx1) / 6.0;
        wvx += -nfx * len;
        wvy += -nfy * len;
        wlen += len;
        wmx += len * 0.5 * (sx1 + sx2);
        wmy += len * 0.5 * (sy1 + sy2);
        c.walls.push_back(amr_wall{k, b0, 0, patches ? patches[b0] : 0, sx1, sy1, sx2, sy2, -nfx, -nfy, len});
        walled = true;
      }
      const double ae = c.axl[static_cast<size_t>(j) * (nx + 1) + i + 1];
      const double aw = c.axl[static_cast<size_t>(j) * (nx + 1) + i];
      const double an = c.ayl[static_cast<size_t>(j + 1) * nx + i];
      const double as = c.ayl[static_cast<size_t>(j) * nx + i];
      if (!walled) {
        if (cfd::hipamr::inside_polygon(body, a.xc[i + ng], a.yc[j + ng])) {
          c.vol[k] = 0.0;
          c.flag[k] = 2.0;
          ++c.nsolid;
        }
        continue;
      }
      // open faces: east +y, north -x, west -y, south +x
      A += x1 * ae - x0 * aw;
      MX += 0.5 * x1 * x1 * ae - 0.5 * x0 * x0 * aw;
      MY += -0.5 * y1 * y1 * (-an) - 0.5 * y0 * y0 * as;
      c.wax[k] = wvx;
      c.way[k] = wvy;
      if (wlen > 0.0) {
        c.wcx[k] = wmx / wlen;
        c.wcy[k] = wmy / wlen;
      }
      if (A < 1e-13 * a.dx[i + ng] * a.dy[j + ng]) {
        c.walls.resize(first_wall);
        c.vol[k] = 0.0;
        c.flag[k] = 2.0;
        ++c.nsolid;
        continue;
      }
      c.vol[k] = A;
      c.cx[k] = MX / A;
      c.cy[k] = MY / A;
      c.flag[k] = 1.0;
      ++c.ncut;
    }

  // 4. faces next to solid cells closed
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const bool lsol = i > 0 && c.vol[static_cast<size_t>(j + ng) * mx + (i - 1 + ng)] == 0.0;
      const bool rsol = i < nx && c.vol[static_cast<size_t>(j + ng) * mx + (i + ng)] == 0.0;
      if (lsol || rsol) c.axl[static_cast<size_t>(j) * (nx + 1) + i] = 0.0;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const bool bsol = j > 0 && c.vol[static_cast<size_t>(j - 1 + ng) * mx + (i + ng)] == 0.0;
      const bool tsol = j < ny && c.vol[static_cast<size_t>(j + ng) * mx + (i + ng)] == 0.0;
      if (bsol || tsol) c.ayl[static_cast<size_t>(j) * nx + i] = 0.0;
    }

  // 5. clearance from the domain boundary
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (c.flag[static_cast<size_t>(j + ng) * mx + (i + ng)] == 0.0) continue;
      if (i < 3 || i >= nx - 3 || j < 3 || j >= ny - 3)
        return report(err, err_len, 2,
                      "cut geometry: body too close to the domain boundary (cell " + std::to_string(i) + "," +
                          std::to_string(j) + ")");
    }

  // 6. SRD neighbourhoods
  std::vector<int> small;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t k = static_cast<size_t>(j + ng) * mx + (i + ng);
      if (c.vol[k] > 0.0 && c.vol[k] < 0.5 * a.dx[i + ng] * a.dy[j + ng]) small.push_back(static_cast<int>(k));
    }
  const int ns = static_cast<int>(small.size());
  c.nsmall = ns;
  c.sidx.assign(ns, 0.0);
  c.scnt.assign(ns, 0.0);
  c.smem.assign(25 * static_cast<size_t>(ns), 0.0);
  c.svhat.assign(ns, 0.0);
  c.sxhat.assign(ns, 0.0);
  c.syhat.assign(ns, 0.0);
  for (int q = 0; q < ns; ++q) {
    c.sidx[q] = small[q];
    c.smallpos[small[q]] = q;
  }
  auto vol_at = [&](int i, int j) { return c.vol[static_cast<size_t>(j + ng) * mx + (i + ng)]; };
  for (int q = 0; q < ns; ++q) {
    const int k = small[q];
    const int i = k % mx - ng, j = k / mx - ng;
    const double half = 0.5 * a.dx[i + ng] * a.dy[j + ng];
    int cnt = 0, mem[25];
    bool enough = false;
    const double wx = c.wax[k], wy = c.way[k];
    const double wl = std::sqrt(wx * wx + wy * wy);
    if (wl > 1e-14) {
      const double fx = -wx / wl, fy = -wy / wl;  // into the fluid
      int di = 0, dj = 0;
      if (std::fabs(fx) >= std::fabs(fy))
        di = fx > 0 ? 1 : -1;
      else
        dj = fy > 0 ? 1 : -1;
      const int ii = i + di, jj = j + dj;
      if (ii >= 0 && ii < nx && jj >= 0 && jj < ny && vol_at(ii, jj) > 0.0 && vol_at(i, j) + vol_at(ii, jj) >= half) {
        mem[cnt++] = k;
        mem[cnt++] = (jj + ng) * mx + (ii + ng);
        enough = true;
      }
    }
    for (int r = 1; r <= 2 && !enough; ++r) {
      cnt = 0;
      double total = 0.0;
      for (int dj = -r; dj <= r; ++dj)
        for (int di = -r; di <= r; ++di) {
          const int ii = i + di, jj = j + dj;
          if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) continue;
          const double v = vol_at(ii, jj);
          if (v <= 0.0) continue;
          mem[cnt++] = (jj + ng) * mx + (ii + ng);
          total += v;
        }
      if (total >= half) enough = true;
    }
    if (!enough) ++c.nbad;
    c.scnt[q] = cnt;
    for (int m = 0; m < cnt; ++m) {
      c.smem[25 * static_cast<size_t>(q) + m] = mem[m];
      if (mem[m] != k) c.nbrcount[mem[m]] += 1.0;
    }
  }
  for (int q = 0; q < ns; ++q) {
    double vh = 0.0, xh = 0.0, yh = 0.0;
    for (int m = 0; m < static_cast<int>(c.scnt[q]); ++m) {
      const size_t r = static_cast<size_t>(c.smem[25 * static_cast<size_t>(q) + m]);
      const double w = c.vol[r] / c.nbrcount[r];
      vh += w;
      xh += w * c.cx[r];
      yh += w * c.cy[r];
    }
    c.svhat[q] = vh;
    c.sxhat[q] = xh / vh;
    c.syhat[q] = yh / vh;
  }
  return 0;
}

}  // namespace

extern "C" int ugrid_cut_build(const ugrid_axes* axes, const double* px, const double* py, int32_t n,
                               const int32_t* patches, ugrid_cut** out, char* err, size_t err_len) {
  if (!out) return report(err, err_len, 1, "cut geometry: null output handle");
  *out = nullptr;
  if (!axes || !px || !py || axes->nx < 1 || axes->ny < 1 || !axes->xf || !axes->xc || !axes->dx || !axes->yf ||
      !axes->yc || !axes->dy)
    return report(err, err_len, 1, "cut geometry: invalid grid");
  std::unique_ptr<ugrid_cut> c(new ugrid_cut());
  const int status = build(*axes, px, py, n, patches, *c, err, err_len);
  if (status) return status;
  *out = c.release();
  return 0;
}

extern "C" int ugrid_cut_get(const ugrid_cut* c, ugrid_cut_view* v) {
  if (!c || !v) return 1;
  v->nsmall = c->nsmall;
  v->nbad_nbhd = c->nbad;
  v->nsolid = c->nsolid;
  v->ncut = c->ncut;
  v->n_walls = c->walls.size();
  v->vol = c->vol.data();
  v->cx = c->cx.data();
  v->cy = c->cy.data();
  v->flag = c->flag.data();
  v->axl = c->axl.data();
  v->axy = c->axy.data();
  v->ayl = c->ayl.data();
  v->ayx = c->ayx.data();
  v->wax = c->wax.data();
  v->way = c->way.data();
  v->wcx = c->wcx.data();
  v->wcy = c->wcy.data();
  v->nbrcount = c->nbrcount.data();
  v->smallpos = c->smallpos.data();
  v->sidx = c->sidx.data();
  v->scnt = c->scnt.data();
  v->smem = c->smem.data();
  v->svhat = c->svhat.data();
  v->sxhat = c->sxhat.data();
  v->syhat = c->syhat.data();
  v->walls = c->walls.data();
  return 0;
}

extern "C" void ugrid_cut_destroy(ugrid_cut* c) { delete c; }
