// Oracle (test infrastructure only) for the guiding network: the MLP, its
// von Mises-Fisher mixture head, the training loss and its gradient, in FP64.
//
// The reference has no path guiding; this restates the network that
// include/lajolla_amd.h specifies (canonical parameter layout, one-blob
// inputs, 4 x 64 ReLU, 8 lobes x (mean, log kappa, logit)).  With
// emulate_bf16 set, the values the device feeds to its BF16 matrix cores
// (weights, one-blob inputs, hidden activations, output and hidden deltas)
// are rounded FP64 -> FP32 -> BF16 (round to nearest even) at the same
// points, so device and oracle differ only by FP32 accumulation order.
//
// Pinned by tests/test_oracle_guide.py: the analytic gradient against central
// finite differences of the FP64 loss; the mixture pdf integrates to one and
// has the closed-form vMF peak; sampled directions have the pdf's moments.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 160, kHid = 64, kOut = 40, kLobes = 8, kLayers = 5;
constexpr int kNumParams = 25384;
constexpr double kPi = 3.14159265358979323846;
constexpr double kLogKappaLo = -3.0, kLogKappaHi = 8.0;

int layer_in(int l) { return l == 0 ? kIn : kHid; }
int layer_out(int l) { return l == 4 ? kOut : kHid; }
int param_w(int l) { return l == 0 ? 0 : param_w(l - 1) + layer_in(l - 1) * layer_out(l - 1) + layer_out(l - 1); }
int param_b(int l) { return param_w(l) + layer_in(l) * layer_out(l); }

double to_bf16(double x) {
    float f = float(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return double(f);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return double(f);
}
double fp32(double x) { return double(float(x)); }

// One-blob code of the 7 features: px, py, pz over 32 bins, the four angle
// features over 16 bins; bin b of feature s is exp(-((b + 1/2) - s nb)^2 / 2).
void encode(const float *feat, double *x) {
    int k = 0;
    for (int f = 0; f < 7; f++) {
        int nb = f < 3 ? 32 : 16;
        for (int b = 0; b < nb; b++) {
            double d = (b + 0.5) - double(feat[f]) * nb;
            x[k++] = std::exp(-0.5 * d * d);
        }
    }
}

struct Net {
    std::vector<double> p;
    bool em;
    double q(double v) const { return em ? to_bf16(v) : v; }
    double w(int l, int i, int o) const { return q(p[param_w(l) + i * layer_out(l) + o]); }
    double b(int l, int o) const { return em ? fp32(p[param_b(l) + o]) : p[param_b(l) + o]; }
};

struct Trace {
    double x[kIn];
    double h[4][kHid];    // ReLU outputs (rounded when emulating)
    double y[kOut];
};

void forward(const Net &N, const float *feat, Trace &T) {
    encode(feat, T.x);
    for (int i = 0; i < kIn; i++) T.x[i] = N.q(T.x[i]);
    const double *in = T.x;
    for (int l = 0; l < kLayers; l++) {
        int ni = layer_in(l), no = layer_out(l);
        for (int o = 0; o < no; o++) {
            double s = 0;
            for (int i = 0; i < ni; i++) s += N.w(l, i, o) * in[i];
            s += N.b(l, o);
            if (l < 4) T.h[l][o] = N.q(std::max(s, 0.0));
            else T.y[o] = s;
        }
        if (l < 4) in = T.h[l];
    }
}

struct Mix {
    double mu[kLobes][3], kappa[kLobes], len[kLobes], logw[kLobes];
    bool inside[kLobes];
};

double log_norm(double k) { return std::log(k) - std::log(2 * kPi) - std::log(-std::expm1(-2 * k)); }

void mixture(const double *y, Mix &m) {
    double lmax = -1e300;
    for (int j = 0; j < kLobes; j++) lmax = std::max(lmax, y[5 * j + 4]);
    double z = 0;
    for (int j = 0; j < kLobes; j++) z += std::exp(y[5 * j + 4] - lmax);
    double log_z = lmax + std::log(z);
    for (int j = 0; j < kLobes; j++) {
        const double *r = y + 5 * j;
        double l2 = std::max(r[0] * r[0] + r[1] * r[1] + r[2] * r[2], 1e-12);
        m.len[j] = std::sqrt(l2);
        for (int c = 0; c < 3; c++) m.mu[j][c] = r[c] / m.len[j];
        m.inside[j] = r[3] > kLogKappaLo && r[3] < kLogKappaHi;
        m.kappa[j] = std::exp(std::min(std::max(r[3], kLogKappaLo), kLogKappaHi));
        m.logw[j] = r[4] - log_z;
    }
}

// log q(dir) and d log q / d y
double log_q_grad(const double *y, const double *d, double *gy) {
    Mix m;
    mixture(y, m);
    double a[kLobes], mw[kLobes], amax = -1e300;
    for (int j = 0; j < kLobes; j++) {
        mw[j] = m.mu[j][0] * d[0] + m.mu[j][1] * d[1] + m.mu[j][2] * d[2];
        a[j] = m.logw[j] + log_norm(m.kappa[j]) + m.kappa[j] * (mw[j] - 1);
        amax = std::max(amax, a[j]);
    }
    double s = 0;
    for (int j = 0; j < kLobes; j++) s += std::exp(a[j] - amax);
    double lq = amax + std::log(s);
    if (gy) {
        for (int j = 0; j < kLobes; j++) {
            double gam = std::exp(a[j] - lq), pi = std::exp(m.logw[j]), k = m.kappa[j];
            gy[5 * j + 4] = gam - pi;
            gy[5 * j + 3] = m.inside[j] ? gam * (1 - 2 * k / std::expm1(2 * k) + k * (mw[j] - 1)) : 0.0;
            double sc = gam * k / m.len[j];
            for (int c = 0; c < 3; c++) gy[5 * j + c] = sc * (d[c] - mw[j] * m.mu[j][c]);
        }
    }
    return lq;
}

}  // namespace

extern "C" {

int orc_guide_num_params(void) { return kNumParams; }

// Raw outputs (n x 40) for feature rows (n x 8 floats).
void orc_guide_forward(const double *params, const float *feats, int n, int emulate_bf16, double *raw_out) {
    Net N{std::vector<double>(params, params + kNumParams), emulate_bf16 != 0};
    Trace T;
    for (int i = 0; i < n; i++) {
        forward(N, feats + 8 * i, T);
        std::memcpy(raw_out + size_t(i) * kOut, T.y, sizeof(T.y));
    }
}

// Mixture pdf of raw outputs at a unit direction.
double orc_guide_pdf(const double *raw, const double *dir) { return std::exp(log_q_grad(raw, dir, nullptr)); }

// Mixture sampling: lobe by its weight with u_lobe, then the vMF inversion
// (Wood 1994) with (u1, u2) in the basis of Duff et al. 2017.
void orc_guide_sample(const double *raw, double u_lobe, double u1, double u2, double *dir) {
    Mix m;
    mixture(raw, m);
    int j = kLobes - 1;
    double c = 0;
    for (int k = 0; k < kLobes; k++) {
        c += std::exp(m.logw[k]);
        if (u_lobe < c) {
            j = k;
            break;
        }
    }
    double k = m.kappa[j];
    double w = 1 + std::log(u1 + (1 - u1) * std::exp(-2 * k)) / k;
    w = std::min(std::max(w, -1.0), 1.0);
    double st = std::sqrt(std::max(0.0, 1 - w * w)), ph = 2 * kPi * u2;
    const double *mu = m.mu[j];
    double sg = std::copysign(1.0, mu[2]);
    double a = -1 / (sg + mu[2]), b = mu[0] * mu[1] * a;
    double t1[3] = {1 + sg * mu[0] * mu[0] * a, sg * b, -sg * mu[0]};
    double t2[3] = {b, sg + mu[1] * mu[1] * a, -mu[1]};
    for (int c2 = 0; c2 < 3; c2++) dir[c2] = w * mu[c2] + st * (std::cos(ph) * t1[c2] + std::sin(ph) * t2[c2]);
}

// Normalised loss -sum w log q / sum w over a batch (dirs: n x 4 = xyz, w)
// and its gradient with respect to the canonical parameters.
double orc_guide_loss_grad(const double *params, const float *feats, const float *dirs, int n, int emulate_bf16,
                           double *grad_out) {
    Net N{std::vector<double>(params, params + kNumParams), emulate_bf16 != 0};
    std::vector<double> g(kNumParams, 0.0);
    double loss = 0, wsum = 0;
    Trace T;
    for (int i = 0; i < n; i++) {
        const float *dd = dirs + 4 * i;
        double d[3] = {dd[0], dd[1], dd[2]}, w = dd[3];
        if (emulate_bf16) {
            // the device evaluates the loss in FP32 on FP32 directions
            for (double &c : d) c = fp32(c);
        }
        forward(N, feats + 8 * i, T);
        double gy[kOut];
        double lq = log_q_grad(T.y, d, gy);
        loss += -w * lq;
        wsum += w;
        // delta at the outputs: d(-w log q)/dy
        double dz[kHid], dh[kHid], prev[kHid];
        double dy[kOut];
        for (int o = 0; o < kOut; o++) dy[o] = N.q(-w * gy[o]);
        // output layer
        for (int o = 0; o < kOut; o++) {
            g[param_b(4) + o] += dy[o];
            for (int k = 0; k < kHid; k++) g[param_w(4) + k * kOut + o] += T.h[3][k] * dy[o];
        }
        const double *delta = dy;
        for (int l = 4; l >= 1; l--) {
            int no = layer_out(l);
            for (int k = 0; k < kHid; k++) {
                double s = 0;
                for (int o = 0; o < no; o++) s += N.w(l, k, o) * delta[o];
                dh[k] = s;
            }
            for (int k = 0; k < kHid; k++) dz[k] = N.q(T.h[l - 1][k] > 0 ? dh[k] : 0.0);
            int lp = l - 1;
            const double *in = lp == 0 ? T.x : T.h[lp - 1];
            int ni = layer_in(lp);
            for (int k = 0; k < kHid; k++) {
                g[param_b(lp) + k] += dz[k];
                for (int i2 = 0; i2 < ni; i2++) g[param_w(lp) + i2 * kHid + k] += in[i2] * dz[k];
            }
            std::memcpy(prev, dz, sizeof(dz));
            delta = prev;
        }
    }
    if (grad_out)
        for (int k = 0; k < kNumParams; k++) grad_out[k] = wsum > 0 ? g[k] / wsum : 0.0;
    return wsum > 0 ? loss / wsum : 0.0;
}

// Features of a surface vertex: position in the box [lo, hi], then normal and
// outgoing direction as (theta / pi, phi / 2 pi).
void orc_guide_features(const double *pos, const double *nrm, const double *wo, const double *lo, const double *hi,
                        float *out8) {
    for (int c = 0; c < 3; c++) {
        double e = hi[c] - lo[c];
        double v = e > 0 ? (pos[c] - lo[c]) / e : 0.5;
        out8[c] = float(std::min(std::max(v, 0.0), 1.0));
    }
    auto ang = [](const double *v, float *o) {
        double th = std::acos(std::min(std::max(v[2], -1.0), 1.0));
        double ph = std::atan2(v[1], v[0]);
        if (ph < 0) ph += 2 * kPi;
        o[0] = float(th / kPi);
        o[1] = float(ph / (2 * kPi));
    };
    ang(nrm, out8 + 3);
    ang(wo, out8 + 5);
    out8[7] = 0.f;
}

}  // extern "C"
