// pam_cli.h -- host side of the C++ drivers run_modeling / run_migration:
// Parameters.json, the geometry CSVs and raw float32 model / data files, and
// the host preprocessing of the C++ reference (wavelet, mute, smoothing),
// feeding libpam through its C ABI (include/pam.h).
//
// Mirrors cpp/Survey.hpp:6-130 / cpp/Survey.cpp:20-176 (the parsed keys and
// the derived sizes with std::round), cpp/Modeling.cu:84-120 (Ricker, Cerjan)
// and cpp/Migration.cu:86-236 (Mute, createMask, smoothModel).
#pragma once
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <limits>
#include <vector>

#include "pam.h"

namespace cli {

const float kPi = 3.14159265358979323846f;

// ------------------------------------------------------------- a small JSON reader
// Enough of JSON for Parameters.json: objects, arrays, strings, numbers,
// true / false / null.  Values are kept as text (strings unquoted) or, for
// arrays, as their elements.
struct JValue {
    std::string text;
    std::vector<JValue> items;
    bool is_array = false;
};

class JsonReader {
   public:
    explicit JsonReader(const std::string &s) : s_(s) {}
    std::map<std::string, JValue> object() {
        std::map<std::string, JValue> out;
        ws();
        expect('{');
        ws();
        if (peek() == '}') {
            ++i_;
            return out;
        }
        for (;;) {
            ws();
            std::string key = string();
            ws();
            expect(':');
            ws();
            out[key] = value();
            ws();
            if (peek() == ',') {
                ++i_;
                continue;
            }
            expect('}');
            return out;
        }
    }

   private:
    const std::string &s_;
    size_t i_ = 0;
    char peek() const { return i_ < s_.size() ? s_[i_] : '\0'; }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    void expect(char c) {
        if (peek() != c) throw std::runtime_error(std::string("Parameters.json: expected '") + c + "'");
        ++i_;
    }
    std::string string() {
        expect('"');
        std::string out;
        while (peek() != '"') {
            if (peek() == '\0') throw std::runtime_error("Parameters.json: unterminated string");
            if (peek() == '\\') ++i_;
            out += s_[i_++];
        }
        ++i_;
        return out;
    }
    JValue value() {
        JValue v;
        if (peek() == '"') {
            v.text = string();
        } else if (peek() == '[') {
            ++i_;
            v.is_array = true;
            ws();
            if (peek() == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                v.items.push_back(value());
                ws();
                if (peek() == ',') {
                    ++i_;
                    continue;
                }
                expect(']');
                break;
            }
        } else if (peek() == '{') {
            (void)object();
        } else {
            while (i_ < s_.size() && s_[i_] != ',' && s_[i_] != '}' && s_[i_] != ']' &&
                   !std::isspace((unsigned char)s_[i_]))
                v.text += s_[i_++];
        }
        return v;
    }
};

inline int cround(float v) { return (int)std::round(v); }

// --------------------------------------------------------------------- the survey
struct Survey {
    std::map<std::string, JValue> p;
    std::string approximation, migration, ABC;
    float dx = 0, dz = 0, dt = 0, L = 0, D = 0, T = 0, fcut = 0, tlag = 0;
    int N_abc = 0, nx = 0, nz = 0, nt = 0, itlag = 0, nt_data = 0, nx_abc = 0, nz_abc = 0;
    float sigma = 0, shift = 0, window = 0, v0 = 1500;
    bool reciprocity = false, mirror = false;
    int idx_water = 0, step = 0, last_save = 0;
    bool snap = false;
    float cerjan_factor = 0;
    std::string rec_file, src_file, vpFile, epsilonFile, deltaFile, thetaFile;
    std::string seismogramFolder, imageFolder, snapshotFolder;
    std::vector<float> rec_x, rec_z, shot_x, shot_z;
    std::vector<int> rx, rz, sx, sz;
    int Nrec = 0, Nshot = 0;

    const JValue &get(const std::string &k) const {
        auto it = p.find(k);
        if (it == p.end()) throw std::runtime_error("Parameters.json: missing key " + k);
        return it->second;
    }
    std::string str(const std::string &k, const std::string &d = "") const {
        auto it = p.find(k);
        return it == p.end() ? d : it->second.text;
    }
    float num(const std::string &k, float d) const {
        auto it = p.find(k);
        return it == p.end() ? d : std::stof(it->second.text);
    }
    bool flag(const std::string &k) const { return str(k, "false") == "true"; }

    explicit Survey(const std::string &path) {
        std::ifstream f(path);
        if (!f.is_open()) throw std::invalid_argument("could not open " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        p = JsonReader(text).object();
        approximation = get("approximation").text;
        migration = str("migration", "onthefly");
        ABC = get("ABC").text;
        dx = std::stof(get("dx").text);
        dz = std::stof(get("dz").text);
        dt = std::stof(get("dt").text);
        L = std::stof(get("L").text);
        D = std::stof(get("D").text);
        T = std::stof(get("T").text);
        N_abc = (int)std::stof(get("N_abc").text);
        fcut = std::stof(get("fcut").text);
        sigma = num("sigma", 0.f);
        shift = num("shift", 0.f);
        window = num("window", 0.f);
        v0 = num("v0", 1500.f);
        step = (int)num("step", 0.f);
        last_save = (int)num("last_save", 0.f);
        snap = flag("snap");
        idx_water = (int)num("idx_water", 0.f);
        cerjan_factor = num("cerjan_factor", 0.f);
        reciprocity = flag("reciprocity");
        mirror = flag("mirror");
        rec_file = str("rec_file");
        src_file = str("src_file");
        vpFile = str("vpFile");
        epsilonFile = str("epsilonFile");
        deltaFile = str("deltaFile");
        thetaFile = str("thetaFile");
        seismogramFolder = str("seismogramFolder");
        imageFolder = str("imageFolder");
        snapshotFolder = str("snapshotFolder");
        // derived sizes, cpp/Survey.cpp:101-108
        tlag = 2.0f * std::sqrt(kPi) / fcut;
        itlag = cround(tlag / dt);
        nt_data = cround(T / dt) + 1;
        nx = cround(L / dx) + 1;
        nz = cround(D / dz) + 1;
        nt = itlag + nt_data;
        nx_abc = nx + 2 * N_abc;
        nz_abc = nz + 2 * N_abc;
        readGeometry();
    }

    static void read_csv(const std::string &path, std::vector<float> &x, std::vector<float> &z) {
        std::ifstream file(path);
        if (!file.is_open()) throw std::invalid_argument("could not open " + path);
        std::string line;
        std::getline(file, line);  // header
        while (std::getline(file, line)) {
            if (line.find_first_not_of(" \r\n\t") == std::string::npos) continue;
            std::stringstream ls(line);
            std::string index, cx, cz;
            std::getline(ls, index, ',');
            std::getline(ls, cx, ',');
            std::getline(ls, cz, ',');
            x.push_back(std::stof(cx));
            z.push_back(std::stof(cz));
        }
    }

    // cpp/Survey.cpp:128-176
    void readGeometry() {
        read_csv(rec_file, rec_x, rec_z);
        read_csv(src_file, shot_x, shot_z);
        if (reciprocity) {
            std::swap(shot_x, rec_x);
            std::swap(shot_z, rec_z);
        }
        Nrec = (int)rec_x.size();
        Nshot = (int)shot_x.size();
        for (int i = 0; i < Nrec; ++i) {
            rx.push_back(cround(rec_x[i] / dx) + N_abc);
            rz.push_back(cround(rec_z[i] / dz) + N_abc - (mirror ? idx_water : 0));
        }
        for (int i = 0; i < Nshot; ++i) {
            sx.push_back(cround(shot_x[i] / dx) + N_abc);
            sz.push_back(cround(shot_z[i] / dz) + N_abc - (mirror ? idx_water : 0));
        }
    }

    int physics() const {
        if (approximation == "acoustic") return PAM_ACOUSTIC;
        if (approximation == "VTI") return PAM_VTI;
        if (approximation == "TTI") return PAM_TTI;
        throw std::invalid_argument("unknown approximation " + approximation);
    }
    int abc() const {
        if (ABC == "cerjan") return PAM_CERJAN;
        if (ABC == "CPML") return PAM_CPML;
        throw std::invalid_argument("unknown ABC " + ABC);
    }
    std::string fcut_text() const {
        std::ostringstream o;
        o << std::fixed << std::setprecision(1) << fcut;
        return o.str();
    }
};

// ------------------------------------------------------------------ file I/O
inline std::vector<float> read_bin(const std::string &path, size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::invalid_argument("could not open " + path);
    std::vector<float> v(n);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(float)));
    if ((size_t)f.gcount() != n * sizeof(float))
        throw std::runtime_error(path + ": short read (" + std::to_string(f.gcount()) + " of " +
                                 std::to_string(n * sizeof(float)) + " bytes)");
    return v;
}
inline void write_bin(const std::string &path, const float *v, size_t n) {
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::invalid_argument("could not open " + path);
    f.write(reinterpret_cast<const char *>(v), (std::streamsize)(n * sizeof(float)));
    std::cout << "Info: File saved to " << path << std::endl;
}

// ---------------------------------------------------------- host preprocessing
// plain Ricker of peak fcut/3 scaled by 1/(dx dz), float32 (cpp/Modeling.cu:84-100)
inline std::vector<float> ricker(const Survey &s) {
    std::vector<float> w(s.nt);
    const float scale = 1.0f / (s.dx * s.dz);
    const float fc = s.fcut / (3.0f * std::sqrt(kPi));
    for (int n = 0; n < s.nt; ++n) {
        const float td = n * s.dt - s.tlag;
        const float arg = kPi * kPi * kPi * fc * fc * td * td;
        w[n] = (1.0f - 2.0f * arg) * std::exp(-arg) * scale;
    }
    return w;
}

// Dispersion and stability check of the C++ driver (cpp/Modeling.cu:184-327),
// float32: acoustic lambda_min / 4.28 and dx_lim / (4 vmax); VTI / TTI the dt
// limit sqrt(2) / (sqrt(S) max(vp sqrt(A))) with S the 8th-order stencil sum
// and A the squared slowness factor at the Nyquist wavenumbers (Sd gated at
// |den| <= 1e-12), and the qP phase velocity over 181 angles in [0, pi/2] (VTI)
// or 361 in [0, pi] relative to the tilt (TTI, theta in radians).  Prints the
// driver's report; the return value says whether dx, dz and dt pass.
struct Stability {
    float vmin = 0, vmax = 0, dx_lim = 0, dt_lim = 0;
    bool ok = false;
};
inline float sd_unit(float e, float d, float s2, float c2) {
    const float num = -2.0f * (e - d) * s2 * c2;
    const float den = (1.0f + 2.0f * e) * (s2 * s2) + (c2 * c2) + 2.0f * (1.0f + d) * s2 * c2;
    return std::fabs(den) > 1e-12f ? num / den : 0.0f;
}
inline Stability check_dispersion_stability(const Survey &s, const std::vector<float> &vp,
                                            const std::vector<float> &eps, const std::vector<float> &del,
                                            const std::vector<float> &th) {
    Stability r;
    float vmin = vp[0], vmax = vp[0];
    for (float v : vp) {
        vmin = std::min(vmin, v);
        vmax = std::max(vmax, v);
    }
    const bool aniso = s.approximation == "VTI" || s.approximation == "TTI";
    float emax = 0.0f, dmax = 0.0f;
    if (!aniso) {
        r.dx_lim = vmin / s.fcut / 4.28f;
        r.dt_lim = r.dx_lim / (4.0f * vmax);
    } else {
        const float c[4] = {8.0f / 5.0f, -1.0f / 5.0f, 8.0f / 315.0f, -1.0f / 560.0f};
        float S = 0.0f;
        for (int m = 1; m <= 4; ++m) S += c[m - 1] * (m % 2 ? 2.0f : 0.0f);
        const float idx2 = (1.0f / s.dx) * (1.0f / s.dx), idz2 = (1.0f / s.dz) * (1.0f / s.dz);
        float smax = 0.0f;
        for (size_t i = 0; i < vp.size(); ++i) {
            emax = std::max(emax, eps[i]);
            dmax = std::max(dmax, del[i]);
            const float Sk = sd_unit(eps[i], del[i], idx2, idz2);
            const float A = ((1.0f + 2.0f * eps[i]) + Sk) * idx2 + (1.0f + Sk) * idz2;
            smax = std::max(smax, vp[i] * std::sqrt(A));
        }
        r.dt_lim = std::sqrt(2.0f) / (std::sqrt(S) * smax);
        const bool tti = s.approximation == "TTI";
        const int nang = tti ? 361 : 181;
        const float amax = tti ? kPi : 0.5f * kPi;
        vmin = std::numeric_limits<float>::infinity();
        vmax = 0.0f;
        for (int ia = 0; ia < nang; ++ia) {
            const float phi = amax * (float)ia / (float)(nang - 1);
            for (size_t i = 0; i < vp.size(); ++i) {
                const float b = tti ? phi - th[i] : phi;
                const float sn = std::sin(b), cs = std::cos(b);
                const float s2 = sn * sn, c2 = cs * cs;
                const float v = vp[i] * std::sqrt((1.0f + 2.0f * eps[i]) * s2 + c2 + sd_unit(eps[i], del[i], s2, c2));
                vmin = std::min(vmin, v);
                vmax = std::max(vmax, v);
            }
        }
        r.dx_lim = vmin / s.fcut / 4.28f;
    }
    r.vmin = vmin;
    r.vmax = vmax;
    r.ok = s.dx <= r.dx_lim && s.dz <= r.dx_lim && s.dt <= r.dt_lim;
    std::cout << std::setprecision(9);
    if (aniso) {
        std::cout << "info: Maximum epsilon: " << emax << "\n";
        std::cout << "info: Maximum delta: " << dmax << "\n";
    }
    std::cout << "info: Dispersion and stability check\n"
              << "info: Minimum velocity: " << vmin << " m/s\n"
              << "info: Maximum velocity: " << vmax << " m/s\n"
              << "info: Maximum frequency: " << s.fcut << " Hz\n"
              << "info: Current dx: " << s.dx << " m\n"
              << "info: Current dz: " << s.dz << " m\n"
              << "info: Current dt: " << s.dt << " s\n"
              << "info: Critical dx for dispersion: " << r.dx_lim << " m\n"
              << "info: Critical dz for dispersion: " << r.dx_lim << " m\n"
              << "info: Critical dt for stability: " << r.dt_lim << " s\n"
              << (r.ok ? "info: Dispersion and stability conditions satisfied."
                       : "WARNING: Dispersion or stability conditions not satisfied.")
              << std::endl;
    std::cout << std::setprecision(6);
    return r;
}

// water mask |f - min f| < 1e-3 (cpp/Migration.cu:131-145)
inline std::vector<char> water_mask(const std::vector<float> &f) {
    float mn = f[0];
    for (float v : f) mn = std::min(mn, v);
    std::vector<char> m(f.size());
    for (size_t i = 0; i < f.size(); ++i) m[i] = std::fabs(f[i] - mn) < 1e-3f;
    return m;
}

// --------------------------------------------------------------- device helpers
#define CLI_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define CLI_PAM(ctx, call)                                                                     \
    do {                                                                                       \
        int r_ = (call);                                                                       \
        if (r_ != PAM_OK) throw std::runtime_error(std::string(#call) + ": " + pam_last_error(ctx)); \
    } while (0)

inline float *to_device(const std::vector<float> &h) {
    float *d = nullptr;
    CLI_HIP(hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(float)));
    CLI_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return d;
}

// separable Gaussian smoothing of slowness (or of a parameter), clamped
// borders, masked cells restored: the C++ driver's rule (cpp/Migration.cu:147-236)
// run on the GPU by pam_smooth_model
inline void smooth(const Survey &s, std::vector<float> &f, const std::vector<char> &mask, bool parameter) {
    const size_t n = f.size();
    float *d_in = to_device(f), *d_out = nullptr;
    unsigned char *d_mask = nullptr;
    const std::vector<unsigned char> m(mask.begin(), mask.end());
    CLI_HIP(hipMalloc(&d_out, n * sizeof(float)));
    CLI_HIP(hipMalloc(&d_mask, n));
    CLI_HIP(hipMemcpy(d_mask, m.data(), n, hipMemcpyHostToDevice));
    if (pam_smooth_model(d_in, d_out, d_mask, s.nz, s.nx, s.sigma, parameter ? 1 : 0, PAM_PREP_CPP, nullptr) != PAM_OK)
        throw std::runtime_error(std::string("pam_smooth_model: ") + pam_last_error(nullptr));
    CLI_HIP(hipMemcpy(f.data(), d_out, n * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_mask);
}

// linear top mute of every shot's gather, device [Nshot][nt_data][Nrec] in
// place: the C++ driver's rule (cpp/Migration.cu:86-115) run on the GPU by pam_mute
inline void mute_all(const Survey &s, float *d_data) {
    const std::vector<double> rx(s.rec_x.begin(), s.rec_x.end()), rz(s.rec_z.begin(), s.rec_z.end());
    const std::vector<double> sx(s.shot_x.begin(), s.shot_x.end()), sz(s.shot_z.begin(), s.shot_z.end());
    if (pam_mute(d_data, s.Nshot, s.nt_data, s.Nrec, rx.data(), rz.data(), sx.data(), sz.data(), s.dt, s.shift,
                 s.window, s.v0, PAM_PREP_CPP, nullptr) != PAM_OK)
        throw std::runtime_error(std::string("pam_mute: ") + pam_last_error(nullptr));
}

// context with the survey's grid, physics, boundary, wavelet and geometry
inline pam_ctx *make_context(const Survey &s) {
    pam_grid g{s.nz, s.nx, s.N_abc, s.dz, s.dx, s.dt, s.nt, s.itlag};
    // CPML constants of src/Modeling2D.py:110-126 (the C++ reference has no CPML)
    double R = std::pow(1e-3, s.N_abc / 10.0);
    if (s.N_abc >= 200) R = std::pow(1e-3, 15.0);
    pam_abc_params ap{s.cerjan_factor, (float)(-3.0 * std::log(R)), s.fcut / 3.0f};
    pam_ctx *ctx = nullptr;
    const int rc = pam_create(&g, s.physics(), s.abc(), &ap, nullptr, std::max(1, s.Nshot), &ctx);
    if (rc != PAM_OK) {
        const std::string msg = ctx ? pam_last_error(ctx) : "pam_create failed";
        if (ctx) pam_destroy(ctx);
        throw std::runtime_error(msg);
    }
    std::vector<float> w = ricker(s);
    float *dw = to_device(w);
    CLI_PAM(ctx, pam_set_wavelet(ctx, dw));
    CLI_HIP(hipFree(dw));
    CLI_PAM(ctx, pam_set_geometry(ctx, s.sz.data(), s.sx.data(), s.Nshot, s.rz.data(), s.rx.data(), s.Nrec));
    return ctx;
}

}  // namespace cli
