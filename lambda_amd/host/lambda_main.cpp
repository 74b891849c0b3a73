// Command-line renderer: `lambda <scene.lrd> [out.ppm]`.
//
// Same flow as the reference's main (src/main.cpp:27-173): validate the
// arguments, parse the scene, render with progress reports every 100 ms,
// write a PPM.  The render runs on the GPU through lambda_hip_render while
// this thread polls the finished-pixel counter.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "lambda_hip.h"
#include "lambda_host.h"

static bool ends_with(const std::string &s, const std::string &e) {
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

int main(int argc, char **argv) {
    if (argc != 2 && argc != 3) {
        std::fprintf(stderr, "%s: invalid number of arguments\n", argv[0]);
        return 1;
    }
    std::string input = argv[1];
    if (!ends_with(input, ".lrd")) {
        std::fprintf(stderr, "%s: invalid file extension\n", input.c_str());
        return 1;
    }
    std::string output;
    if (argc == 3) {
        output = argv[2];
        if (!ends_with(output, ".ppm")) {
            std::fprintf(stderr, "%s: invalid file extension\n", output.c_str());
            return 1;
        }
    } else {
        output = input.substr(0, input.size() - 4) + ".ppm";
    }

    char err[4096];
    lambda_scene *scene = nullptr;
    if (lambda_scene_load(input.c_str(), &scene, err, sizeof err) != LAMBDA_OK) {
        std::fprintf(stderr, "%s\n", err);
        return 1;
    }
    lambda_scene_view view;
    lambda_render_params params;
    lambda_scene_get_view(scene, &view);
    lambda_scene_get_params(scene, &params);
    const int total = params.width * params.height;
    std::vector<float> rgb(size_t(total) * 3);

    volatile int32_t progress = 0;
    std::atomic<int> rc{LAMBDA_OK};
    std::printf("\r%s: 0%% complete", argv[0]);
    std::fflush(stdout);
    std::thread worker([&] {
        rc = lambda_hip_render(&view, &params, rgb.data(), &progress, err, sizeof err);
    });
    int done = 0;
    while (done < total && rc == LAMBDA_OK) {
        done = progress;
        std::printf("\r%s: %d%% complete", argv[0], int(100.0 * done / total));
        std::fflush(stdout);
        if (done >= total) break;
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
    }
    worker.join();
    if (rc != LAMBDA_OK) {
        std::fprintf(stderr, "\n%s: %s\n", argv[0], err);
        lambda_scene_free(scene);
        return 1;
    }
    std::printf("\r%s: 100%% complete\n", argv[0]);
    int w = params.width, h = params.height;
    lambda_scene_free(scene);
    if (lambda_write_ppm(output.c_str(), w, h, rgb.data(), err, sizeof err) != LAMBDA_OK) {
        std::fprintf(stderr, "%s\n", err);
        return 1;
    }
    return 0;
}
