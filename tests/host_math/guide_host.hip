// Test driver: the guiding distribution of guide.h (encoding, vMF mixture
// density, sampling, loss gradient) compiled for the host.
//
// argv: encode   <features.f32 [n][6]>            <out.f32 [n][96]>
//       encode2  <features.f32 [n][6]>            <out.f32 [n][96]>  (by halves)
//       pdf      <outputs.f32 [n][40]> <dirs.f32 [n][3]> <out.f32 [n]>
//       pdfc     (as pdf, through the compact form)
//       grad     <outputs.f32 [n][40]> <dirs.f32 [n][3]> <c.f32 [n]> <out.f32 [n][41] = loss, d[40]>
//       sample   <outputs.f32 [40]> <count> <seed>   <out.f32 [count][3]>
//       samplec  (as sample, through the compact form)
//       table    <geometry.f32 [7] = lo xyz, inv_cell, nx, ny, nz> <points.f32 [n][6] = p xyz, normal xyz>
//                <out.f32 [n][7] = entry, then the entry's centre xyz and normal xyz>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "guide.h"

using namespace lam;

static std::vector<float> read_f32(const char *path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    size_t n = size_t(f.tellg());
    f.seekg(0);
    std::vector<float> v(n / 4);
    f.read(reinterpret_cast<char *>(v.data()), std::streamsize(n));
    return v;
}

static void write_f32(const char *path, const std::vector<float> &v) {
    std::ofstream o(path, std::ios::binary);
    o.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * 4));
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const std::string mode = argv[1];
    if (mode == "encode") {
        std::vector<float> f = read_f32(argv[2]);
        const size_t n = f.size() / kGuideScalars;
        std::vector<float> out(n * kGuideIn);
        // the kernels' encoding: one sincos per scalar, double-angle bands
        for (size_t r = 0; r < n; r++)
            for (int s = 0; s < kGuideScalars; s++) {
                float v[16];
                guide_encode_scalar(f[r * kGuideScalars + s], v);
                for (int k = 0; k < 16; k++) out[r * kGuideIn + s * 16 + k] = v[k];
            }
        write_f32(argv[3], out);
    } else if (mode == "encode2") {
        // the inference kernel's encoding: each half of a scalar's bands
        // from its own sincos
        std::vector<float> f = read_f32(argv[2]);
        const size_t n = f.size() / kGuideScalars;
        std::vector<float> out(n * kGuideIn);
        for (size_t r = 0; r < n; r++)
            for (int s = 0; s < kGuideScalars; s++)
                for (int h = 0; h < 2; h++) {
                    float v[8];
                    guide_encode_half(f[r * kGuideScalars + s], h, v);
                    for (int k = 0; k < 8; k++) out[r * kGuideIn + s * 16 + 8 * h + k] = v[k];
                }
        write_f32(argv[3], out);
    } else if ((mode == "pdf" || mode == "pdfc") && argc >= 5) {
        std::vector<float> o = read_f32(argv[2]), w = read_f32(argv[3]);
        const size_t n = w.size() / 3;
        std::vector<float> out(n);
        for (size_t r = 0; r < n; r++) {
            const V3 dir = v3(w[3 * r], w[3 * r + 1], w[3 * r + 2]);
            if (mode == "pdf") {
                Lobe L[kLobes];
                guide_lobes(&o[r * kLobeOut], L);
                out[r] = guide_pdf(L, dir);
            } else {
                alignas(16) float c[kLobeOut];
                guide_compact(&o[r * kLobeOut], c);
                out[r] = guide_pdf_c(c, dir);
            }
        }
        write_f32(argv[4], out);
    } else if (mode == "grad" && argc >= 6) {
        std::vector<float> o = read_f32(argv[2]), w = read_f32(argv[3]), c = read_f32(argv[4]);
        const size_t n = c.size();
        std::vector<float> out(n * (kLobeOut + 1));
        for (size_t r = 0; r < n; r++) {
            float d[kLobeOut];
            out[r * (kLobeOut + 1)] = guide_loss_grad(&o[r * kLobeOut], v3(w[3 * r], w[3 * r + 1], w[3 * r + 2]), c[r], d);
            std::memcpy(&out[r * (kLobeOut + 1) + 1], d, sizeof d);
        }
        write_f32(argv[5], out);
    } else if ((mode == "sample" || mode == "samplec") && argc >= 6) {
        std::vector<float> o = read_f32(argv[2]);
        const int count = std::atoi(argv[3]);
        Pcg rng = pcg_seed(uint64_t(std::atoll(argv[4])), 7);
        Lobe L[kLobes];
        guide_lobes(o.data(), L);
        alignas(16) float c[kLobeOut];
        guide_compact(o.data(), c);
        std::vector<float> out(size_t(count) * 3);
        for (int i = 0; i < count; i++) {
            float u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
            V3 d = mode == "sample" ? guide_sample(L, u0, u1, u2) : guide_sample_c(c, u0, u1, u2);
            out[3 * size_t(i)] = d.x;
            out[3 * size_t(i) + 1] = d.y;
            out[3 * size_t(i) + 2] = d.z;
        }
        write_f32(argv[5], out);
    } else if (mode == "table" && argc >= 5) {
        std::vector<float> g = read_f32(argv[2]), pts = read_f32(argv[3]);
        GuideTableDev T{};
        for (int a = 0; a < 3; a++) {
            T.lo[a] = g[size_t(a)];
            T.n[a] = int(g[size_t(4 + a)]);
        }
        T.inv_cell = g[3];
        const size_t n = pts.size() / 6;
        std::vector<float> out(n * 7);
        for (size_t i = 0; i < n; i++) {
            const float *q = &pts[i * 6];
            const long long e = guide_table_entry(T, v3(q[0], q[1], q[2]), v3(q[3], q[4], q[5]));
            V3 c, nn;
            guide_table_query(T, e, c, nn);
            float *o = &out[i * 7];
            o[0] = float(e);
            o[1] = c.x;
            o[2] = c.y;
            o[3] = c.z;
            o[4] = nn.x;
            o[5] = nn.y;
            o[6] = nn.z;
        }
        write_f32(argv[4], out);
    } else {
        return 2;
    }
    return 0;
}
