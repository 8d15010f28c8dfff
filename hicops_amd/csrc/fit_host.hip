// Host FP64 tail fit for the spectra whose device fit is too close to a
// boundary to be trusted.
//
// expeRT::ModelTailFit (expeRT.cpp:572-718) truncates slope and intercept
// x 1e6 to ints after a log10 of the survival function, and picks the fit
// window by comparing log10 values with thresholds.  The device's log10 and
// the host libm's (the one the reference runs on) are each correct to a few
// ulp but not identical, so where a value x 1e6 lies within their error
// bound of an integer, or a log10 value within it of a window threshold,
// the device lists the spectrum (HCX_RES_FIT_HOST) and the fit is decided
// here, serially, with the same operations in the same order as
// tailfit_wave in k_query.hip and std::log10.
#include <algorithm>
#include <cmath>

#include "hcx_internal.h"

namespace hcx {

void host_tailfit(const uint32_t *H, uint32_t hbins, uint32_t cpsms, int32_t hyp, double *sx, int32_t out4[4],
                  uint32_t *flags) {
    out4[0] = out4[1] = out4[2] = out4[3] = 0;
    *flags = 0;
    if (hyp < 1 || hyp >= (int)hbins) {
        *flags = HCX_RES_FIT_UB;
        return;
    }
    // end1: last p <= hyp-1 with yy[p] >= 1, else hyp-1; stt1: first p in
    // [0, end1] with yy[p] >= 1, else 0
    int end1 = hyp - 1;
    for (int p = hyp - 1; p >= 0; p--)
        if (H[p] >= 1u) {
            end1 = p;
            break;
        }
    int stt1 = 0;
    for (int p = 0; p <= end1; p++)
        if (H[p] >= 1u) {
            stt1 = p;
            break;
        }
    if (stt1 == end1) end1 += 1;
    const int n = end1 - stt1 + 1;
    out4[2] = stt1;
    out4[3] = end1;
    double mu_t = 0.0, beta_t = 4.0;
    uint32_t fl = 0;
    if (hyp < 1 || n < 2 || end1 >= (int)hbins) {
        fl |= HCX_RES_FIT_UB;
        mu_t = 0.0;
        beta_t = 0.0;
    } else if ((int)cpsms < 1) {
        mu_t = 0.0;
        beta_t = 100.0;
    } else {
        const double vaa = (double)(int)cpsms;
        double run = 0.0;   // integer prefix sums: exact
        for (int i = 0; i < n; i++) {
            run += (double)H[stt1 + i];
            double t = run / vaa;
            t = t + (-1.0);
            sx[i] = -t;
        }
        for (int i = 0; i < n; i++)
            if (sx[i] > 1.0) sx[i] = 0.999;
        int rep = n - 1;
        for (int p = n - 1; p >= 0; p--)
            if (sx[p] >= 1e-4) {
                rep = p;
                break;
            }
        const double rv = sx[rep];
        for (int i = 0; i < n; i++) {
            double v = sx[i];
            if (v <= 0.0) v = rv;
            sx[i] = std::log10(v);
        }
        const double s0 = sx[0];
        const double hgt = sx[n - 1] - s0;
        auto largmax = [&](double v) {
            for (int i = 0; i < n; i++)
                if (sx[i] <= v) return i;
            return 0;
        };
        auto rargmax = [&](double v) {
            for (int p = n - 1; p >= 0; p--)
                if (sx[p] >= v) return p;
            return n - 1;
        };
        int mark = 0, mark2 = 0;
        if (n > 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = rargmax(s0 + hgt * 0.87);
            if (mark2 == n) mark2 -= 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else if (n == 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = n - 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else {
            mark = 0;
            mark2 = n - 1;
        }
        if (mark < 0) {
            fl |= HCX_RES_FIT_UB;
            mu_t = 0.0;
            beta_t = 0.0;
        } else {
            const int m = mark2 - mark + 1;
            const double *y = sx + mark;
            if (m == 1) {
                mu_t = 0.0;
                beta_t = y[0];
            } else {
                double xbar = 0.0, ybar = 0.0, top = 0.0, bot = 0.0;
                for (int i = 0; i < m; i++) {
                    xbar = xbar + (double)(stt1 + mark + i);
                    ybar = ybar + y[i];
                }
                xbar = xbar / (double)m;
                ybar = ybar / (double)m;
                for (int i = 0; i < m; i++) {
                    const double dx = (double)(stt1 + mark + i) - xbar;
                    top = top + dx * (y[i] - ybar);
                    bot = bot + dx * dx;
                }
                mu_t = top / bot;
                beta_t = ybar - mu_t * xbar;
            }
            if (!(std::fabs(mu_t) < 2.0e3) || !(std::fabs(beta_t) < 2.0e3)) {
                fl |= HCX_RES_FIT_UB;
                mu_t = 0.0;
                beta_t = 0.0;
            }
        }
    }
    out4[0] = (int32_t)(mu_t * 1e6);
    out4[1] = (int32_t)(beta_t * 1e6);
    *flags = fl;
}

double host_tailfit_merged(const double *pd, uint32_t hbins, int32_t pN, int32_t hyp, double *sx) {
    if (hyp < 1 || hyp >= (int)hbins) return INFINITY;
    int end1 = hyp - 1;
    for (int p = hyp - 1; p >= 0; p--)
        if (pd[p] >= 1.0) {
            end1 = p;
            break;
        }
    int stt1 = 0;
    for (int p = 0; p <= end1; p++)
        if (pd[p] >= 1.0) {
            stt1 = p;
            break;
        }
    const int n = end1 - stt1 + 1;
    const int vaa = pN;
    double mu_t = 0.0, beta_t = 4.0;
    if (vaa < 1) {
        mu_t = 0.0;
        beta_t = 100.0;
    } else {
        int acc = 1;
        for (int i = 0; i < n; i++) acc = (int)((double)acc + pd[stt1 + i]);
        const int vaa2 = acc > vaa ? acc : vaa;
        double run = 0.0;
        for (int i = 0; i < n; i++) {
            const double q = pd[stt1 + i] / (double)vaa2;
            run = (i == 0) ? q : run + q;
            const double t = run + (-1.0);
            sx[i] = -t;
        }
        for (int i = 0; i < n; i++)
            if (sx[i] > 1.0) sx[i] = 0.999;
        int rep = n - 1;
        for (int p = n - 1; p >= 0; p--)
            if (sx[p] >= 1e-4) {
                rep = p;
                break;
            }
        const double rv = sx[rep];
        for (int i = 0; i < n; i++) {
            double v = sx[i];
            if (v <= 0.0) v = rv;
            sx[i] = std::log10(v);
        }
        const double s0 = sx[0];
        const double hgt = sx[n - 1] - s0;
        auto largmax = [&](double v) {
            for (int i = 0; i < n; i++)
                if (sx[i] <= v) return i;
            return 0;
        };
        auto rargmax = [&](double v) {
            for (int p = n - 1; p >= 0; p--)
                if (sx[p] >= v) return p;
            return n - 1;
        };
        int mark = 0, mark2 = 0;
        if (n > 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = rargmax(s0 + hgt * 0.87);
            if (mark2 == n) mark2 -= 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else if (n == 3) {
            mark = largmax(s0 + hgt * 0.22) - 1;
            mark2 = n - 1;
            if (mark >= mark2) mark = mark2 - 1;
        } else {
            mark = 0;
            mark2 = n - 1;
        }
        if (mark < 0) {
            mu_t = 0.0;
            beta_t = 0.0;
        } else {
            const int m = mark2 - mark + 1;
            const double *y = sx + mark;
            if (m == 1) {
                mu_t = 0.0;
                beta_t = y[0];
            } else {
                double xbar = 0.0, ybar = 0.0, top = 0.0, bot = 0.0;
                for (int i = 0; i < m; i++) {
                    xbar = xbar + (double)(stt1 + mark + i);
                    ybar = ybar + y[i];
                }
                xbar = xbar / (double)m;
                ybar = ybar / (double)m;
                for (int i = 0; i < m; i++) {
                    const double dx = (double)(stt1 + mark + i) - xbar;
                    top = top + dx * (y[i] - ybar);
                    bot = bot + dx * dx;
                }
                mu_t = top / bot;
                beta_t = ybar - mu_t * xbar;
            }
        }
    }
    return std::pow(10, hyp * mu_t + beta_t) * vaa;
}

int fix_host_fits(hcx_ctx *c, const uint32_t *fix, hcx_result *res, const uint32_t *hist, uint32_t hbins,
                  double expect_max) {
    if (!c->fix_host || !c->fix_dev.p || hbins > c->hbins) return set_error("host fit: staging not set up");
    uint32_t *hcnt = (uint32_t *)c->fix_host;
    uint32_t *hrows = (uint32_t *)((char *)c->fix_host + 256);
    hcx_result *hres = (hcx_result *)(hrows + (size_t)kFixRows * c->hbins);
    double *hsx = (double *)(hres + kFixRows);
    uint32_t *drows = (uint32_t *)c->fix_dev.p;
    hcx_result *dres = (hcx_result *)(drows + (size_t)kFixRows * c->hbins);
    HCX_HIP(hipMemcpyAsync(hcnt, fix, 4, hipMemcpyDeviceToHost, c->stream));
    HCX_HIP(hipStreamSynchronize(c->stream));
    const uint32_t cnt = *hcnt;
    for (uint32_t base = 0; base < cnt; base += kFixRows) {
        const uint32_t k = std::min(kFixRows, cnt - base);
        const uint32_t *ids = fix + 1 + base;
        int rc = launch_fix_gather(c, ids, k, hist, hbins, res, drows, dres);
        if (rc) return rc;
        HCX_HIP(hipMemcpyAsync(hrows, drows, 4ull * k * hbins, hipMemcpyDeviceToHost, c->stream));
        HCX_HIP(hipMemcpyAsync(hres, dres, sizeof(hcx_result) * k, hipMemcpyDeviceToHost, c->stream));
        HCX_HIP(hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < k; i++) {
            hcx_result &r = hres[i];
            int32_t o4[4];
            uint32_t fl = 0;
            // the device stored the hyperscore's histogram bin in maxhypscore
            host_tailfit(hrows + (size_t)i * hbins, hbins, r.cpsms, r.maxhypscore, hsx, o4, &fl);
            r.mu = o4[0];
            r.beta = o4[1];
            r.minhypscore = o4[2];
            r.nexthypscore = o4[3];
            double w_ = r.mu, b_ = r.beta;
            w_ /= 1e6;
            b_ /= 1e6;
            const double lgs = (w_ * r.maxhypscore) + b_;
            double e = std::pow(10.0, lgs);
            e *= (double)r.cpsms;
            r.e_x = e;
            uint16_t flags = r.flags & ~(HCX_RES_FIT_UB | HCX_RES_REPORTED);
            flags |= (uint16_t)fl | HCX_RES_FIT_HOST;
            if (e < expect_max) flags |= HCX_RES_REPORTED;
            r.flags = flags;
        }
        HCX_HIP(hipMemcpyAsync(dres, hres, sizeof(hcx_result) * k, hipMemcpyHostToDevice, c->stream));
        rc = launch_fix_scatter(c, ids, k, dres, res);
        if (rc) return rc;
        HCX_HIP(hipStreamSynchronize(c->stream));
    }
    return HCX_SUCCESS;
}

}  // namespace hcx
