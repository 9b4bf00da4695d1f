// TEST INFRASTRUCTURE ONLY — FP64 CPU reference of the pole-pair overlap quantity and its
// gradient (the product's definition in include/jagua_amd.h, jg_overlap_*), computed in
// double precision from the same f32 inputs.  Used as the checker of the GPU forward and
// backward kernels, together with central finite differences of this same function.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// pose [n][3] (x, y, theta), poles [total][3] (px, py, r), off [n+1], pairs [m][2],
// cedges [k][4].  Writes Q to *q and, when grad is non-null, dQ/d(x, y, theta) [n][3].
void or_overlap64(const double* pose, const float* poles, const uint32_t* off, uint32_t n, const uint32_t* pairs,
                  uint64_t m, const float* cedges, uint32_t k, double* q, double* grad) {
    auto world = [&](uint32_t item, uint32_t i, double& cx, double& cy, double& rx, double& ry) {
        const double x = pose[3 * item], y = pose[3 * item + 1], t = pose[3 * item + 2];
        const double s = std::sin(t), c = std::cos(t);
        const double px = poles[3 * i], py = poles[3 * i + 1];
        rx = c * px - s * py;
        ry = s * px + c * py;
        cx = rx + x;
        cy = ry + y;
    };
    double Q = 0.0;
    if (grad) std::memset(grad, 0, sizeof(double) * 3 * n);
    for (uint64_t p = 0; p < m; p++) {
        const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        for (uint32_t i = off[a]; i < off[a + 1]; i++) {
            double cix, ciy, rix, riy;
            world(a, i, cix, ciy, rix, riy);
            for (uint32_t j = off[b]; j < off[b + 1]; j++) {
                double cjx, cjy, rjx, rjy;
                world(b, j, cjx, cjy, rjx, rjy);
                const double dx = cix - cjx, dy = ciy - cjy;
                const double sq = dx * dx + dy * dy;
                const double rs = (double)poles[3 * i + 2] + (double)poles[3 * j + 2];
                if (sq < rs * rs) {
                    const double d = std::sqrt(sq);
                    const double pen = rs - d;
                    Q += pen * pen;
                    if (grad && d > 0) {
                        const double kk = -2.0 * pen / d;
                        const double gx = kk * dx, gy = kk * dy;
                        grad[3 * a] += gx;
                        grad[3 * a + 1] += gy;
                        grad[3 * a + 2] += -gx * riy + gy * rix;
                        grad[3 * b] -= gx;
                        grad[3 * b + 1] -= gy;
                        grad[3 * b + 2] += gx * rjy - gy * rjx;
                    }
                }
            }
        }
    }
    // container term: the pole's signed distance D to the container boundary (the k edges,
    // positive inside by ray-crossing parity), penetration r - D when positive
    for (uint32_t a = 0; a < n && k > 0; a++)
        for (uint32_t i = off[a]; i < off[a + 1]; i++) {
            double cx, cy, rx, ry;
            world(a, i, cx, cy, rx, ry);
            const double r = poles[3 * i + 2];
            double best = INFINITY, bx = 0.0, by = 0.0;
            uint32_t crossings = 0;
            for (uint32_t e = 0; e < k; e++) {
                const double x1 = cedges[4 * e], y1 = cedges[4 * e + 1], x2 = cedges[4 * e + 2], y2 = cedges[4 * e + 3];
                const double ex = x2 - x1, ey = y2 - y1;
                const double l2 = ex * ex + ey * ey;
                double t = l2 != 0 ? ((cx - x1) * ex + (cy - y1) * ey) / l2 : -1.0;
                double qx, qy;
                if (t < 0) { qx = x1; qy = y1; }
                else if (t > 1) { qx = x2; qy = y2; }
                else { qx = x1 + t * ex; qy = y1 + t * ey; }
                const double dx = cx - qx, dy = cy - qy;
                const double d = std::sqrt(dx * dx + dy * dy);
                if (d < best) { best = d; bx = dx; by = dy; }
                if ((y1 > cy) != (y2 > cy) && cx < x1 + (cy - y1) * ex / ey) crossings++;
            }
            const double sgn = (crossings & 1u) ? 1.0 : -1.0;
            const double pen = r - sgn * best;
            if (pen > 0) {
                Q += pen * pen;
                if (grad && best > 0) {
                    const double kk = -2.0 * pen * sgn / best;
                    grad[3 * a] += kk * bx;
                    grad[3 * a + 1] += kk * by;
                    grad[3 * a + 2] += -kk * bx * ry + kk * by * rx;
                }
            }
        }
    *q = Q;
}

}  // extern "C"
