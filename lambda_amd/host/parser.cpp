// .lrd scene parser and .obj mesh loader of the host front end.
//
// Grammar and behaviour follow Lambda's Parser (core/parser.h:29-786, grammar
// in README.md:20-82): one command per line, '#' comments, brackets padded
// into tokens and commas treated as blanks; the same validation messages
// ("<file>:<line>: <message>"); spectra resampled onto the 1 nm grid with
// linear interpolation in double precision.  Token extraction goes through
// std::istringstream exactly as the reference does, so numbers parse to the
// same floats and edge cases (partially numeric tokens, missing fields) end
// in the same errors.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "scene_model.h"

namespace lambda_host {

namespace {

std::string lowercase_first(std::string s) {
    if (!s.empty()) s[0] = char(std::tolower(s[0]));
    return s;
}

std::runtime_error open_error(const std::string &path) {
    return std::runtime_error(path + ": failed to open file (" + lowercase_first(std::strerror(errno)) + ")");
}

std::string missing(const std::string &param) { return "'" + param + "' is missing or invalid"; }

// Comment stripping and bracket padding (core/parser.h:384-401).
std::string preprocess(const std::string &line) {
    size_t hash = line.find('#');
    std::string body = hash == std::string::npos ? line : line.substr(0, hash);
    std::string out;
    out.reserve(body.size() + 16);
    for (char c : body) {
        if (c == '(' || c == ')' || c == '[' || c == ']') {
            out += ' ';
            out += c;
            out += ' ';
        } else if (c == ',') {
            out += ' ';
        } else {
            out += c;
        }
    }
    return out;
}

// ------------------------------------------------------------------- OBJ

struct ObjData {
    std::vector<Vec3> v, vn;
    std::vector<float> tu, tv;
};

class ObjLoader {
  public:
    ObjLoader(const std::string &path, std::vector<Object> &objects, int material)
        : path_(path), objects_(objects), material_(material) {}

    void run() {
        std::ifstream in(path_);
        if (!in.is_open()) throw open_error(path_);
        std::string line;
        bool saw_face = false;
        while (std::getline(in, line)) {
            line_++;
            if (line.empty()) continue;
            if (!fast_line(line, saw_face)) exact_line(line, saw_face);
        }
        if (!saw_face) throw err("expected \"f\"");
    }

  private:
    std::string path_;
    std::vector<Object> &objects_;
    int material_;
    int line_ = 0;
    ObjData d_;

    std::runtime_error err(const std::string &m) const {
        return std::runtime_error(path_ + ":" + std::to_string(line_) + ": " + m);
    }

    static bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

    // A plain decimal number that strtof consumes completely; anything else
    // (and any value istream extraction might treat differently) takes the
    // exact path.
    static bool plain_float(const char *b, const char *e, float &out) {
        if (b == e || e - b > 64) return false;
        for (const char *p = b; p < e; p++) {
            char c = *p;
            if (!((c >= '0' && c <= '9') || c == '.' || c == '-' || c == '+' || c == 'e' || c == 'E')) return false;
        }
        char buf[72];
        std::memcpy(buf, b, size_t(e - b));
        buf[e - b] = 0;
        char *end = nullptr;
        errno = 0;
        float v = std::strtof(buf, &end);
        if (end != buf + (e - b) || errno != 0 || !std::isfinite(v)) return false;
        if (v != 0 && std::fabs(v) < 1.17549435e-38f) return false;
        out = v;
        return true;
    }

    static bool plain_int(const char *b, const char *e, long &out) {
        if (b == e || e - b > 10) return false;
        const char *p = b;
        if (*p == '-') p++;
        if (p == e) return false;
        for (const char *q = p; q < e; q++)
            if (*q < '0' || *q > '9') return false;
        out = std::strtol(std::string(b, e).c_str(), nullptr, 10);
        return true;
    }

    // Resolves a 1-based (or negative, relative) OBJ index
    // (core/parser.h:716-751).  The reference leaves index 0 undefined; it is
    // rejected here.
    int resolve(long raw, size_t size, const char *what) const {
        long n = long(size);
        if (raw < 0 && n + raw >= 0) return int(n + raw);
        if (raw <= 0) throw err(std::string("'") + what + "' must be positive, got " + std::to_string(raw));
        if (raw > n) throw err(std::string("'") + what + "' must be at most " + std::to_string(n - 1) + ", got " + std::to_string(raw));
        return int(raw - 1);
    }

    void emit_face(const std::vector<int> &vi, const std::vector<int> &ti, const std::vector<int> &ni) {
        if (vi.size() < 3) throw err("'f' must have at least 3 vertices");
        Vec3 p0 = d_.v[vi[0]];
        Vec3 n0 = ni[0] >= 0 ? d_.vn[ni[0]] : Vec3(0, 0, 0);
        float u0 = ti[0] >= 0 ? d_.tu[ti[0]] : 0;
        float v0 = ti[0] >= 0 ? d_.tv[ti[0]] : 0;
        // Fan triangulation around the first corner (core/parser.h:761-779).
        for (size_t i = 1; i + 1 < vi.size(); i++) {
            Vec3 p1 = d_.v[vi[i]];
            Vec3 n1 = ni[i] >= 0 ? d_.vn[ni[i]] : Vec3(0, 0, 0);
            float u1 = ti[i] >= 0 ? d_.tu[ti[i]] : 1;
            float v1 = ti[i] >= 0 ? d_.tv[ti[i]] : 0;
            Vec3 p2 = d_.v[vi[i + 1]];
            Vec3 n2 = ni[i + 1] >= 0 ? d_.vn[ni[i + 1]] : Vec3(0, 0, 0);
            float u2 = ti[i + 1] >= 0 ? d_.tu[ti[i + 1]] : 0;
            float v2 = ti[i + 1] >= 0 ? d_.tv[ti[i + 1]] : 1;
            objects_.push_back(Object::tri(material_, p0, p1 - p0, p2 - p0, n0, n1, n2, u0, u1, u2, v0, v1, v2));
        }
    }

    // Fast path for well-formed lines.  Returns false (having changed
    // nothing) when the line needs the exact stream-based path.
    bool fast_line(const std::string &line, bool &saw_face) {
        const char *s = line.data(), *end = s + line.size();
        const char *tb[64], *te[64];
        int nt = 0;
        for (const char *p = s; p < end;) {
            while (p < end && is_blank(*p)) p++;
            if (p >= end) break;
            if (nt == 64) return false;
            tb[nt] = p;
            while (p < end && !is_blank(*p)) p++;
            te[nt++] = p;
        }
        if (nt == 0) return true;  // blank line: the reference skips it too
        std::string tag(tb[0], te[0]);
        if (tag == "v" || tag == "vn") {
            if (nt != 4) return false;
            float x, y, z;
            if (!plain_float(tb[1], te[1], x) || !plain_float(tb[2], te[2], y) || !plain_float(tb[3], te[3], z)) return false;
            (tag == "v" ? d_.v : d_.vn).push_back(Vec3(x, y, z));
            return true;
        }
        if (tag == "vt") {
            if (nt != 3) return false;
            float a, b;
            if (!plain_float(tb[1], te[1], a) || !plain_float(tb[2], te[2], b)) return false;
            d_.tu.push_back(a);
            d_.tv.push_back(b);
            return true;
        }
        if (tag == "f") {
            std::vector<int> vi, ti, ni;
            struct Raw { long v, t, n; bool ht, hn; };
            std::vector<Raw> raws;
            for (int k = 1; k < nt; k++) {
                const char *b = tb[k], *e = te[k];
                const char *s1 = static_cast<const char *>(std::memchr(b, '/', size_t(e - b)));
                Raw r{0, 0, 0, false, false};
                if (!s1) {
                    if (!plain_int(b, e, r.v)) return false;
                } else {
                    if (!plain_int(b, s1, r.v)) return false;
                    const char *b2 = s1 + 1;
                    const char *s2 = static_cast<const char *>(std::memchr(b2, '/', size_t(e - b2)));
                    const char *e2 = s2 ? s2 : e;
                    if (e2 > b2) {
                        if (!plain_int(b2, e2, r.t)) return false;
                        r.ht = true;
                    }
                    if (s2) {
                        const char *b3 = s2 + 1;
                        if (std::memchr(b3, '/', size_t(e - b3))) return false;
                        if (e > b3) {
                            if (!plain_int(b3, e, r.n)) return false;
                            r.hn = true;
                        }
                    }
                }
                raws.push_back(r);
            }
            saw_face = true;
            for (const Raw &r : raws) {
                vi.push_back(resolve(r.v, d_.v.size(), "v"));
                ti.push_back(r.ht ? resolve(r.t, d_.tu.size(), "vt") : -1);
                ni.push_back(r.hn ? resolve(r.n, d_.vn.size(), "vn") : -1);
            }
            emit_face(vi, ti, ni);
            return true;
        }
        return false;
    }

    // Exact path: the reference's stream extraction, line by line
    // (core/parser.h:655-783).
    void exact_line(const std::string &line, bool &saw_face) {
        std::stringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) return;
        if (tag == "v" || tag == "vn") {
            Vec3 p;
            if (!(ss >> p.x)) throw err("'x' is missing or invalid");
            if (!(ss >> p.y)) throw err("'y' is missing or invalid");
            if (!(ss >> p.z)) throw err("'z' is missing or invalid");
            (tag == "v" ? d_.v : d_.vn).push_back(p);
        } else if (tag == "vt") {
            float a, b;
            if (!(ss >> a)) throw err("'u' is missing or invalid");
            if (!(ss >> b)) throw err("'v' is missing or invalid");
            d_.tu.push_back(a);
            d_.tv.push_back(b);
        } else if (tag == "f") {
            saw_face = true;
            std::vector<int> vi, ti, ni;
            std::string tok;
            while (ss >> tok) {
                std::stringstream ts(tok);
                std::string sv, st, sn;
                if (!std::getline(ts, sv, '/')) throw err("'v' is missing or invalid");
                std::getline(ts, st, '/');
                std::getline(ts, sn, '/');
                long raw;
                try { raw = std::stoi(sv); } catch (const std::exception &) { throw err("'v' is missing or invalid"); }
                vi.push_back(resolve(raw, d_.v.size(), "v"));
                if (!st.empty()) {
                    try { raw = std::stoi(st); } catch (const std::exception &) { throw err("'vt' is missing or invalid"); }
                    ti.push_back(resolve(raw, d_.tu.size(), "vt"));
                } else {
                    ti.push_back(-1);
                }
                if (!sn.empty()) {
                    try { raw = std::stoi(sn); } catch (const std::exception &) { throw err("'vn' is missing or invalid"); }
                    ni.push_back(resolve(raw, d_.vn.size(), "vn"));
                } else {
                    ni.push_back(-1);
                }
            }
            emit_face(vi, ti, ni);
        } else {
            return;  // unknown tags are ignored without the trailing-token check
        }
        if (ss >> tag) throw err("unexpected token \"" + tag + "\"");
    }
};

// ------------------------------------------------------------------- LRD

class LrdParser {
  public:
    LrdParser(const std::string &name, Scene &s) : name_(name), s_(s) {}

    void run(std::istream &in) {
        // Slot 0 of every table is the "default" entry (core/parser.h:51-60).
        s_.spectra.push_back(filled(0.5f));
        s_.spectrum_textures.push_back(SpectrumTexture{LAMBDA_SPECTRUM_CONSTANT, 0, 0, 0});
        spectrum_tex_["default"] = 0;
        ScalarTexture def;
        def.type = LAMBDA_SCALAR_CONSTANT;
        def.p[0] = 0.5f;
        s_.scalar_textures.push_back(def);
        scalar_tex_["default"] = 0;
        Material lamb;
        lamb.type = LAMBDA_MATERIAL_LAMBERTIAN;
        s_.materials.push_back(lamb);
        material_["default"] = 0;

        std::string raw;
        while (std::getline(in, raw)) {
            line_++;
            std::string line = preprocess(raw);
            if (line.empty()) continue;
            std::stringstream ss(line);
            std::string cmd;
            if (!(ss >> cmd)) continue;
            if (!have_render_ && cmd != "Render") throw err("expected \"Render\", got \"" + cmd + "\"");
            if (cmd == "Render") render(ss);
            else if (cmd == "Texture") texture(ss);
            else if (cmd == "Material") material(ss);
            else if (cmd == "Object") object(ss);
            else if (cmd == "Instance") instance(ss);
            else if (cmd == "Background") background(ss);
            else if (cmd == "Camera") camera(ss);
            else
                throw err("expected \"Background\", \"Camera\", \"Instance\", \"Material\", \"Render\", \"Object\", or \"Texture\", got \"" + cmd + "\"");
            if (ss >> cmd) throw err("unexpected token \"" + cmd + "\"");
        }
        if (!have_render_) throw std::runtime_error(name_ + ": expected \"Render\"");
        if (!have_background_) throw std::runtime_error(name_ + ": expected \"Background\"");
        if (!have_camera_) throw std::runtime_error(name_ + ": expected \"Camera\"");
        if (!have_instance_) throw std::runtime_error(name_ + ": expected \"Instance\"");

        offsets_.push_back(int(s_.objects.size()));
        build_bvh(s_, offsets_);
        build_lights(s_);
    }

  private:
    std::string name_;
    Scene &s_;
    int line_ = 0;
    bool have_render_ = false, have_background_ = false, have_camera_ = false, have_instance_ = false;
    std::vector<int> offsets_;
    std::map<std::string, int> object_, material_, scalar_tex_, spectrum_tex_;

    std::runtime_error err(const std::string &m) const {
        return std::runtime_error(name_ + ":" + std::to_string(line_) + ": " + m);
    }

    static Spectrum filled(float v) {
        Spectrum s;
        s.fill(v);
        return s;
    }

    static float to_float(const std::string &t, size_t *idx = nullptr) { return std::stof(t, idx); }

    template <typename T>
    T value(std::stringstream &ss, const std::string &param) {
        T v;
        if (!(ss >> v)) throw err(missing(param));
        return v;
    }

    Vec3 vector(std::stringstream &ss, const std::string &param) {
        std::string tok;
        if (!(ss >> tok) || tok != "(") throw err(missing(param));
        Vec3 v;
        if (!(ss >> v.x)) throw err(missing(param));
        if (!(ss >> v.y)) throw err(missing(param));
        if (!(ss >> v.z)) throw err(missing(param));
        if (!(ss >> tok) || tok != ")") throw err(missing(param));
        return v;
    }

    // Complex literal: "a", "bi", "a+bi", "a-i", "i", "-i"; parts are read
    // in float precision (core/parser.h:414-447).  `param` is the caller's
    // full message, which the reference wraps a second time.
    std::complex<double> complex_literal(std::string tok, const std::string &param) {
        std::string m = missing(param);
        double re = 0, im = 0;
        if (tok == "i") return {0, 1};
        if (tok == "-i") return {0, -1};
        try {
            size_t used = 0;
            re = to_float(tok, &used);
            if (used < tok.length()) {
                tok = tok.substr(used);
                if (tok.back() == 'i') tok.pop_back();
                else throw err(m);
                if (tok == "+") im = 1;
                else if (tok == "-") im = -1;
                else if (tok.empty()) {
                    im = re;
                    re = 0;
                } else {
                    im = to_float(tok);
                }
            }
        } catch (const std::exception &) {
            throw err(m);
        }
        return {re, im};
    }

    // A single value or "( lambda value ... )", resampled to 1 nm
    // (core/parser.h:466-561).  Returns the index of the new spectrum.
    template <bool kComplex>
    int spectrum(std::stringstream &ss, const std::string &param) {
        using T = std::conditional_t<kComplex, std::complex<double>, float>;
        std::string m = missing(param);
        std::string tok;
        if (!(ss >> tok)) throw err(m);
        if (tok != "(") {
            if constexpr (!kComplex) {
                float v;
                try { v = to_float(tok); } catch (const std::exception &) { throw err(m); }
                s_.spectra.push_back(filled(v));
                return int(s_.spectra.size()) - 1;
            } else {
                ComplexSpectrum c;
                c.fill(complex_literal(tok, m));
                s_.complex_spectra.push_back(c);
                return int(s_.complex_spectra.size()) - 1;
            }
        }
        std::map<double, T> samples;
        while (ss >> tok) {
            if (tok == ")") break;
            double lambda;
            try { lambda = std::stod(tok); } catch (const std::exception &) { throw err(m); }
            if (lambda < kLambdaMin || lambda > kLambdaMax) throw err(m);
            if (!(ss >> tok)) throw err(m);
            if constexpr (!kComplex) {
                try { samples[lambda] = to_float(tok); } catch (const std::exception &) { throw err(m); }
            } else {
                samples[lambda] = complex_literal(tok, m);
            }
        }
        if (tok != ")" || samples.empty()) throw err(m);

        std::array<T, kBins> out;
        auto it = samples.begin();
        double lo = it->first, hi = samples.rbegin()->first;
        for (int i = 0; i < kBins; i++) {
            if (i <= lo - kLambdaMin || samples.size() == 1) {
                out[i] = samples[lo];
            } else if (i >= hi - kLambdaMin) {
                out[i] = samples[hi];
            } else {
                double lambda = i + kLambdaMin;
                while (std::next(it) != samples.end() && std::next(it)->first < lambda) ++it;
                double a = it->first, b = std::next(it)->first;
                out[i] = T((b - lambda) / (b - a) * samples[a] + (lambda - a) / (b - a) * samples[b]);
            }
        }
        if constexpr (!kComplex) {
            s_.spectra.push_back(out);
            return int(s_.spectra.size()) - 1;
        } else {
            s_.complex_spectra.push_back(out);
            return int(s_.complex_spectra.size()) - 1;
        }
    }

    std::vector<std::string> string_array(std::stringstream &ss, const std::string &param) {
        std::string tok;
        if (!(ss >> tok) || tok != "[") throw err(missing(param));
        std::vector<std::string> out;
        while (ss >> tok) {
            if (tok == "]") return out;
            out.push_back(tok);
        }
        throw err(missing(param));
    }

    std::vector<int> complex_spectrum_array(std::stringstream &ss, const std::string &param) {
        std::string tok;
        if (!(ss >> tok) || tok != "[") throw err(missing(param));
        std::streampos pos = ss.tellg();
        std::vector<int> out;
        while (ss >> tok) {
            if (tok == "]") return out;
            ss.seekg(pos);
            out.push_back(spectrum<true>(ss, param));
            pos = ss.tellg();
        }
        throw err(missing(param));
    }

    // A spectrum literal, or the name of a spectrum or scalar texture
    // (core/parser.h:617-637).
    int spectrum_texture(std::stringstream &ss, const std::string &param) {
        std::streampos pos = ss.tellg();
        try {
            int sp = spectrum<false>(ss, param);
            s_.spectrum_textures.push_back(SpectrumTexture{LAMBDA_SPECTRUM_CONSTANT, sp, 0, 0});
            return int(s_.spectrum_textures.size()) - 1;
        } catch (const std::exception &) {
            ss.seekg(pos);
            std::string name = value<std::string>(ss, param);
            if (spectrum_tex_.count(name)) return spectrum_tex_.at(name);
            if (scalar_tex_.count(name)) {
                s_.spectrum_textures.push_back(SpectrumTexture{LAMBDA_SPECTRUM_SCALAR, scalar_tex_.at(name), 0, 0});
                return int(s_.spectrum_textures.size()) - 1;
            }
            throw err("'" + param + "' is not defined");
        }
    }

    void check_plane(const Vec3 &h, const Vec3 &v) {
        if (length(h) < kEpsilon) throw err("'horizontal' must be non-zero");
        if (length(v) < kEpsilon) throw err("'vertical' must be non-zero");
        if (length(cross(h, v)) < kEpsilon) throw err("'horizontal' and 'vertical' must be non-parallel");
    }

    void render(std::stringstream &ss) {
        have_render_ = true;
        int w = value<int>(ss, "width");
        if (w <= 0) throw err("'width' must be positive, got " + std::to_string(w));
        int h = value<int>(ss, "height");
        if (h <= 0) throw err("'height' must be positive, got " + std::to_string(h));
        int depth = value<int>(ss, "depth");
        if (depth < 0) throw err("'depth' must be non-negative, got " + std::to_string(depth));
        int spp = value<int>(ss, "samples");
        if (spp < 1) throw err("'samples' must be at least 1, got " + std::to_string(spp));
        int root = int(std::sqrt(spp));
        if (root * root != spp) throw err("'samples' must be a perfect square, got " + std::to_string(spp));
        float lmin = value<float>(ss, "lambdaMin");
        if (lmin < 0) throw err("'lambdaMin' must be non-negative, got " + std::to_string(lmin));
        if (lmin < kLambdaMin) throw err("'lambdaMin' must be at least " + std::to_string(kLambdaMin) + ", got " + std::to_string(lmin));
        float lmax = value<float>(ss, "lambdaMax");
        if (lmax < 0) throw err("'lambdaMax' must be non-negative, got " + std::to_string(lmax));
        if (lmax > kLambdaMax) throw err("'lambdaMax' must be at most " + std::to_string(kLambdaMax) + ", got " + std::to_string(lmax));
        if (lmin >= lmax) throw err("'lambdaMin' must be at most 'lambdaMax', got " + std::to_string(lmin) + " and " + std::to_string(lmax));
        s_.width = w;
        s_.height = h;
        s_.depth = depth;
        s_.samples = spp;
        s_.sqrt_samples = root;
        s_.lambda_min = lmin;
        s_.lambda_max = lmax;
        s_.camera = Camera();  // a Render command resets the renderer, camera included
    }

    void texture(std::stringstream &ss) {
        std::string name = value<std::string>(ss, "name");
        std::string type = value<std::string>(ss, "type");
        std::string sub = value<std::string>(ss, "subtype");
        if (scalar_tex_.count(name) || spectrum_tex_.count(name)) throw err("'name' is already defined");
        if (type == "scalar") {
            ScalarTexture t;
            if (sub == "constant") {
                t.type = LAMBDA_SCALAR_CONSTANT;
                t.p[0] = value<float>(ss, "value");
            } else if (sub == "perlin" || sub == "worley") {
                float lo = value<float>(ss, "min");
                float hi = value<float>(ss, "max");
                float f = value<float>(ss, "frequency");
                if (lo > hi) throw err("'min' must be at most 'max', got " + std::to_string(lo) + " and " + std::to_string(hi));
                t.type = sub == "perlin" ? LAMBDA_SCALAR_PERLIN : LAMBDA_SCALAR_WORLEY;
                t.p[0] = lo;
                t.p[1] = hi;
                t.p[2] = f;
            } else {
                throw err("'subtype' must be \"constant\", \"perlin\", or \"worley\"");
            }
            s_.scalar_textures.push_back(t);
            scalar_tex_[name] = int(s_.scalar_textures.size()) - 1;
        } else if (type == "spectrum") {
            if (sub == "constant") {
                int sp = spectrum<false>(ss, "spectrum");
                s_.spectrum_textures.push_back(SpectrumTexture{LAMBDA_SPECTRUM_CONSTANT, sp, 0, 0});
            } else if (sub == "checker") {
                int a = spectrum<false>(ss, "value1");
                int b = spectrum<false>(ss, "value2");
                float scale = value<float>(ss, "scale");
                s_.spectrum_textures.push_back(SpectrumTexture{LAMBDA_SPECTRUM_CHECKER, a, b, scale});
            } else if (sub == "scalar") {
                // The reference reads the texture name twice: once to test it
                // and once more to look it up (core/parser.h:141).
                std::string probe = value<std::string>(ss, "texture");
                if (!scalar_tex_.count(probe)) throw err("'texture' is not defined");
                std::string looked_up = value<std::string>(ss, "texture");
                int idx = scalar_tex_.at(looked_up);
                s_.spectrum_textures.push_back(SpectrumTexture{LAMBDA_SPECTRUM_SCALAR, idx, 0, 0});
            } else {
                throw err("'subtype' must be \"constant\", \"checker\", or \"scalar\"");
            }
            spectrum_tex_[name] = int(s_.spectrum_textures.size()) - 1;
        } else {
            throw err("'type' must be \"scalar\" or \"spectrum\"");
        }
    }

    void material(std::stringstream &ss) {
        std::string name = value<std::string>(ss, "name");
        std::string type = value<std::string>(ss, "type");
        if (material_.count(name)) throw err("'name' is already defined");
        Material m;
        if (type == "lambertian" || type == "mirror") {
            int albedo = spectrum_texture(ss, "albedo");
            float lo = spectrum_texture_min(s_, albedo);
            float hi = spectrum_texture_max(s_, albedo);
            if (lo < 0) throw err("'albedo' must be positive, got " + std::to_string(lo));
            else if (hi > 1) throw err("'albedo' must be at most 1, got " + std::to_string(hi));
            m.type = type == "lambertian" ? LAMBDA_MATERIAL_LAMBERTIAN : LAMBDA_MATERIAL_MIRROR;
            m.p[0] = albedo;
        } else if (type == "dielectric") {
            int n0 = spectrum<false>(ss, "n0");
            int n1 = spectrum<false>(ss, "n1");
            for (int i = 0; i < kBins; i++) {
                if (s_.spectra[n0][i] <= 0) throw err("'n0' must be positive, got " + std::to_string(s_.spectra[n0][i]));
                if (s_.spectra[n1][i] <= 0) throw err("'n1' must be positive, got " + std::to_string(s_.spectra[n1][i]));
            }
            m.type = LAMBDA_MATERIAL_DIELECTRIC;
            m.p[0] = n0;
            m.p[1] = n1;
        } else if (type == "emissive") {
            m.type = LAMBDA_MATERIAL_EMISSIVE;
            m.p[0] = spectrum_texture(ss, "emission");
        } else if (type == "thinfilm") {
            std::vector<int> n = complex_spectrum_array(ss, "n");
            if (n.size() < 2) throw err("'n' must have at least 2 entries, got " + std::to_string(n.size()));
            int n_offset = int(s_.material_properties.size());
            s_.material_properties.insert(s_.material_properties.end(), n.begin(), n.end());
            int layers = int(n.size()) - 2;
            std::vector<std::string> d = string_array(ss, "d");
            if (int(d.size()) != layers) throw err("'d' must have " + std::to_string(layers) + " entries, got " + std::to_string(d.size()));
            int d_offset = int(s_.material_properties.size());
            for (int i = 0; i < layers; i++) {
                // A number becomes a new constant texture; otherwise a name.
                try {
                    float thickness = to_float(d[i]);
                    ScalarTexture t;
                    t.type = LAMBDA_SCALAR_CONSTANT;
                    t.p[0] = thickness;
                    s_.scalar_textures.push_back(t);
                    s_.material_properties.push_back(int(s_.scalar_textures.size()) - 1);
                } catch (const std::exception &) {
                    if (scalar_tex_.count(d[i])) s_.material_properties.push_back(scalar_tex_.at(d[i]));
                    else throw err("'d' is not defined");
                }
            }
            m.type = LAMBDA_MATERIAL_THINFILM;
            m.p[0] = layers;
            m.p[1] = n_offset;
            m.p[2] = d_offset;
        } else {
            throw err("'type' must be \"dielectric\", \"emissive\", \"lambertian\", \"mirror\", or \"thinfilm\"");
        }
        s_.materials.push_back(m);
        material_[name] = int(s_.materials.size()) - 1;
    }

    void object(std::stringstream &ss) {
        std::string name = value<std::string>(ss, "name");
        std::string type = value<std::string>(ss, "type");
        std::string mat = value<std::string>(ss, "material");
        // Consecutive declarations with one name form one object group.
        if (name == "default") throw err("'name' is already defined");
        auto it = object_.find(name);
        if (it != object_.end()) {
            if (it->second != int(offsets_.size()) - 1) throw err("'name' is already defined");
        } else {
            object_[name] = int(offsets_.size());
            offsets_.push_back(int(s_.objects.size()));
        }
        auto mt = material_.find(mat);
        if (mt == material_.end()) throw err("'material' is not defined");
        int material = mt->second;
        if (type == "sphere") {
            Vec3 c = vector(ss, "center");
            float r = value<float>(ss, "radius");
            if (r <= 0) throw err("'radius' must be positive, got " + std::to_string(r));
            s_.objects.push_back(Object::sphere(material, c, r));
        } else if (type == "quad" || type == "tri") {
            Vec3 c = vector(ss, "corner");
            Vec3 h = vector(ss, "horizontal");
            Vec3 v = vector(ss, "vertical");
            check_plane(h, v);
            s_.objects.push_back(type == "quad" ? Object::quad(material, c, h, v) : Object::tri(material, c, h, v));
        } else if (type == "mesh") {
            std::string file = value<std::string>(ss, "file");
            ObjLoader(file, s_.objects, material).run();
        } else {
            throw err("'type' must be \"quad\", \"sphere\", or \"tri\"");
        }
    }

    void instance(std::stringstream &ss) {
        have_instance_ = true;
        std::string obj = value<std::string>(ss, "object");
        std::string mat = value<std::string>(ss, "material");
        auto ot = object_.find(obj);
        if (ot == object_.end()) throw err("'object' is not defined");
        int group = ot->second;
        int first = offsets_[group];
        int count = group < int(offsets_.size()) - 1 ? offsets_[group + 1] - first : int(s_.objects.size()) - first;
        int material = -1;
        if (mat != "default") {
            auto mt = material_.find(mat);
            if (mt == material_.end()) throw err("'material' is not defined");
            material = mt->second;
        }
        Vec3 t = vector(ss, "translation");
        Vec3 r = vector(ss, "rotation");
        Vec3 sc = vector(ss, "scale");
        if (std::fabs(sc.x) < kEpsilon || std::fabs(sc.y) < kEpsilon || std::fabs(sc.z) < kEpsilon)
            throw err("'scale' must be non-zero");
        Instance inst(first, count, material, t, r, sc);
        if (std::fabs(sc.x - sc.y) > kEpsilon || std::fabs(sc.y - sc.z) > kEpsilon)
            for (int i = 0; i < count; i++)
                if (s_.materials[inst.material_of(s_.objects, i)].emissive())
                    throw err("'scale' must be uniform for emissive objects");
        s_.instances.push_back(inst);
    }

    void background(std::stringstream &ss) {
        have_background_ = true;
        std::string type = value<std::string>(ss, "type");
        if (type != "equirectangular") throw err("'type' must be \"equirectangular\"");
        s_.background_texture = spectrum_texture(ss, "background");
    }

    void camera(std::stringstream &ss) {
        have_camera_ = true;
        Camera c;
        c.position = vector(ss, "position");
        c.corner = vector(ss, "corner");
        c.horizontal = vector(ss, "horizontal");
        c.vertical = vector(ss, "vertical");
        check_plane(c.horizontal, c.vertical);
        s_.camera = c;
    }
};

}  // namespace

void parse_lrd_file(const std::string &path, Scene &s) {
    std::ifstream in(path);
    if (!in.is_open()) throw open_error(path);
    LrdParser(path, s).run(in);
}

void parse_lrd_text(const std::string &text, const std::string &name, Scene &s) {
    std::istringstream in(text);
    LrdParser(name, s).run(in);
}

}  // namespace lambda_host
