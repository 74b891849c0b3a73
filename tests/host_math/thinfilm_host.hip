// Test driver: runs the integrator's thin-film code (dscene.h) on the CPU.
// argv[1]: stack = {L, (L+2) x (re, im) indices, L thicknesses}, float32
// argv[2]: queries [n][3] = {cos_theta, lambda, front}, float32
// argv[3]: output reflectances [n], float32
#include <cstdio>
#include <fstream>
#include <vector>

#include "dscene.h"

using namespace lam;

static std::vector<float> read_floats(const char *path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    size_t n = size_t(f.tellg());
    f.seekg(0);
    std::vector<float> v(n / 4);
    f.read(reinterpret_cast<char *>(v.data()), std::streamsize(n));
    return v;
}

int main(int argc, char **argv) {
    if (argc != 4) return 2;
    std::vector<float> st = read_floats(argv[1]), q = read_floats(argv[2]);
    const int L = int(st[0]);
    std::vector<float2> cs(size_t(L + 2) * kBins);
    std::vector<int> props(size_t(2 * L + 2));
    for (int j = 0; j < L + 2; j++) {
        for (int i = 0; i < kBins; i++) cs[size_t(j) * kBins + i] = make_float2(st[1 + 2 * j], st[2 + 2 * j]);
        props[j] = j;
    }
    std::vector<float> th(size_t(L) + 1);
    for (int j = 0; j < L; j++) th[j] = st[1 + 2 * (L + 2) + j];
    DeviceScene S{};
    S.cspectra = cs.data();
    S.props = props.data();
    const int4 mt = make_int4(LAMBDA_MATERIAL_THINFILM, L, 0, L + 2);
    const size_t n = q.size() / 3;
    std::vector<float> out(n);
    for (size_t i = 0; i < n; i++) {
        float c = q[3 * i], lam = q[3 * i + 1];
        bool front = q[3 * i + 2] != 0;
        float s = sqrtf(fmaxf(0.f, 1 - c * c));
        out[i] = thinfilm_reflectance(S, mt, front, v3(0, 0, 1), v3(s, 0, -c), lam, th.data());
    }
    std::ofstream o(argv[3], std::ios::binary);
    o.write(reinterpret_cast<const char *>(out.data()), std::streamsize(n * 4));
    return 0;
}
