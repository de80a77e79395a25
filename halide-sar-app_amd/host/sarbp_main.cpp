// sarbp — SAR backprojection driver with the reference's command line (sarbp.cpp:53-94) and flow
// (sarbp.cpp:96-337): load the platform directory, create the image plane, backproject, convert to
// dB, map to u8 and write a PNG. Every stage runs through the library's C ABI: the backprojection
// and image output on the GPU, the loader and PNG writer on the host.
#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "sarbp_hip.h"
#include "sarbp_host.h"

using namespace std::chrono;

namespace {

// Defaults of sarbp.cpp:44-51 and ImgPlane.h:34-36.
constexpr const char *kSchedDefault = "cpu";
constexpr double kDbMinDefault = 0;
constexpr double kDbMaxDefault = 0;
constexpr int kUpsampleDefault = 6;
constexpr int kTaylorDefault = 20;
constexpr double kResFactorDefault = 1.0;
constexpr double kAspectDefault = 1.0;
const float kNHat[3] = {0.f, 0.f, 1.f};

const char kShortOptions[] = "p:o:d:D:s:t:u:r:a:n:h";
const struct option kLongOptions[] = {
    {"platform-dir", required_argument, nullptr, 'p'},
    {"output", required_argument, nullptr, 'o'},
    {"db-min", required_argument, nullptr, 'd'},
    {"db-max", required_argument, nullptr, 'D'},
    {"schedule", required_argument, nullptr, 's'},
    {"taylor", required_argument, nullptr, 't'},
    {"upsample", required_argument, nullptr, 'u'},
    {"res-factor", required_argument, nullptr, 'r'},
    {"aspect", required_argument, nullptr, 'a'},
    {"num-iters", required_argument, nullptr, 'n'},
    {"help", no_argument, nullptr, 'h'},
    {nullptr, 0, nullptr, 0}};

void print_usage(const std::string &prog, std::ostream &os) {
    os << "Usage: " << prog << " -p DIR -o FILE [OPTION]...\n"
       << "Options:\n"
       << "  -p, --platform-dir=DIR  Platform input directory\n"
       << "  -o, --output=FILE       Output image file (PNG)\n"
       << "  -d, --db-min=REAL       Output image min dB\n"
       << "                          Default: " << kDbMinDefault << "\n"
       << "  -D, --db-max=REAL       Output image max dB\n"
       << "                          Default: " << kDbMaxDefault << "\n"
       << "  -s, --schedule=NAME     One of: cpu, cuda (both run the HIP implementation)\n"
       << "                          Default: " << kSchedDefault << "\n"
       << "  -t, --taylor=INT        Taylor count\n"
       << "                          Default: " << kTaylorDefault << "\n"
       << "  -u, --upsample=INT      Upsample factor\n"
       << "                          Default: " << kUpsampleDefault << "\n"
       << "  -r, --res-factor=REAL   Image res in units of theoretical resolution size\n"
       << "                          Default: " << kResFactorDefault << "\n"
       << "  -a, --aspect=REAL       Aspect ratio of range to cross range\n"
       << "                          Default: " << kAspectDefault << "\n"
       << "  -n, --num-iters=INT     Number of backprojection iterations\n"
       << "                          Default: 1\n"
       << "  -h, --help              Print this message and exit" << std::endl;
}

// A dense Halide buffer over host memory; dims are (extent, stride) from dimension 0 up.
struct HBuf {
    halide_buffer_t b{};
    std::vector<halide_dimension_t> d;
    HBuf(void *host, uint8_t code, uint8_t bits, std::vector<std::pair<int32_t, int32_t>> dims) {
        for (auto &e : dims) d.push_back({0, e.first, e.second, 0});
        b.host = (uint8_t *)host;
        b.type.code = code;
        b.type.bits = bits;
        b.type.lanes = 1;
        b.dimensions = (int32_t)d.size();
        b.dim = d.data();
    }
    halide_buffer_t *operator&() { return &b; }
};

long ms_since(steady_clock::time_point t) {
    return (long)duration_cast<milliseconds>(steady_clock::now() - t).count();
}

}  // namespace

int main(int argc, char **argv) {
    std::string platform_dir, output_png, sched = kSchedDefault;
    int taylor = kTaylorDefault, upsample = kUpsampleDefault;
    double dB_min = kDbMinDefault, dB_max = kDbMaxDefault;
    double res_factor = kResFactorDefault, aspect = kAspectDefault;
    long num_iters = 1;
    for (;;) {
        const int c = getopt_long(argc, argv, kShortOptions, kLongOptions, nullptr);
        if (c == -1) break;
        switch (c) {
        case 'p': platform_dir = optarg; break;
        case 'o': output_png = optarg; break;
        case 'd': dB_min = std::atof(optarg); break;
        case 'D': dB_max = std::atof(optarg); break;
        case 's': sched = optarg; break;
        case 't': taylor = std::atoi(optarg); break;
        case 'u': upsample = std::atoi(optarg); break;
        case 'r': res_factor = std::atof(optarg); break;
        case 'a': aspect = std::atof(optarg); break;
        case 'n': num_iters = std::atol(optarg); break;
        case 'h': print_usage(argv[0], std::cout); return 0;
        default: print_usage(argv[0], std::cerr); return EXIT_FAILURE;
        }
    }
    // constraints of sarbp.cpp:150-181
    auto reject = [&](const char *msg) {
        std::cerr << msg << std::endl;
        print_usage(argv[0], std::cerr);
        return EXIT_FAILURE;
    };
    if (platform_dir.empty() || output_png.empty()) return reject("Missing required parameters");
    if (dB_min >= dB_max) return reject("Constraint failed: dB_min < dB_max");
    if (taylor < 1) return reject("Constraint failed: taylor >= 1");
    if (upsample < 1) return reject("Constraint failed: upsample >= 1");
    if (res_factor <= 0) return reject("Constraint failed: res-factor > 0");
    if (aspect <= 0) return reject("Constraint failed: aspect > 0");
    if (num_iters < 0) return reject("Constraint failed: num-iters >= 0");

    // schedule selection (sarbp.cpp:183-229): both single-node names run the HIP pipeline
    auto bp = backprojection;
    if (sched == "cpu" || sched == "cuda") {
        bp = sched == "cuda" ? backprojection_cuda : backprojection;
        std::cout << "Using schedule " << sched << " (HIP implementation)" << std::endl;
    } else if (sched == "cpu_distributed" || sched == "cuda_distributed") {
        std::cerr << "Distributed schedules require distributed support in Halide" << std::endl;
        return EXIT_FAILURE;
    } else if (sched == "ritsar" || sched == "ritsar-s" || sched == "ritsar-p" ||
               sched == "ritsar-vp" || sched == "auto-m16") {
        std::cerr << "Schedule " << sched << " is not part of this library" << std::endl;
        return EXIT_FAILURE;
    } else {
        std::cerr << "Unknown schedule: " << sched << std::endl;
        print_usage(argv[0], std::cerr);
        return EXIT_FAILURE;
    }

    auto start = steady_clock::now();
    sarbp_platform *pd = nullptr;
    if (sarbp_platform_load(platform_dir.c_str(), &pd) != 0) {
        std::cerr << "sarbp: " << sarbp_host_last_error() << std::endl;
        return EXIT_FAILURE;
    }
    struct Free {
        sarbp_platform *p;
        ~Free() { sarbp_platform_free(p); }
    } free_pd{pd};
    std::cout << "Loaded platform data in " << ms_since(start) << " ms" << std::endl;
    std::cout << "Number of pulses: " << pd->npulses << std::endl;
    std::cout << "Pulse sample size: " << pd->nsamples << std::endl;

    // image plane (ImgPlane.cpp:16-63; sarbp.cpp:248-250 passes the int upsample as a bool)
    const float *n_hat = pd->n_hat ? pd->n_hat : kNHat;
    start = steady_clock::now();
    int32_t dims[2];
    double scal[2];
    int rv = sarbp_img_plane_create(pd->nsamples, pd->npulses, pd->delta_r, pd->R_c, n_hat,
                                    res_factor, aspect, upsample != 0, dims, scal, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr, nullptr);
    const int nu = dims[0], nv = dims[1];
    std::vector<double> u(nu), v(nv), k_u(nu), k_v(nv), u_hat(3), v_hat(3);
    std::vector<double> pixel_locs;
    if (rv == 0) {
        pixel_locs.resize((size_t)3 * (size_t)nu * (size_t)nv);
        rv = sarbp_img_plane_create(pd->nsamples, pd->npulses, pd->delta_r, pd->R_c, n_hat,
                                    res_factor, aspect, upsample != 0, dims, scal, u.data(),
                                    v.data(), k_u.data(), k_v.data(), u_hat.data(), v_hat.data(),
                                    pixel_locs.data());
    }
    if (rv != 0) {
        std::cerr << "sarbp: image plane failed: " << sarbp_last_error() << std::endl;
        return rv;
    }
    std::cout << "Computed image plane parameters in " << ms_since(start) << " ms" << std::endl;
    std::cout << "X length: " << nu << std::endl;
    std::cout << "Y length: " << nv << std::endl;

    // FFT width (sarbp.cpp:258)
    const int N_fft = (int)std::pow(2, (int)std::log2(pd->nsamples * upsample) + 1);
    dft_init_fftw((size_t)N_fft);

    const int ns = pd->nsamples, np = pd->npulses;
    const int32_t npix = nu * nv;
    HBuf b_phs((void *)pd->phs, halide_type_float, 32, {{2, 1}, {ns, 2}, {np, 2 * ns}});
    HBuf b_k_r((void *)pd->k_r, halide_type_float, 32, {{ns, 1}});
    HBuf b_u(u.data(), halide_type_float, 64, {{nu, 1}});
    HBuf b_v(v.data(), halide_type_float, 64, {{nv, 1}});
    HBuf b_pos((void *)pd->pos, halide_type_float, 32, {{3, 1}, {np, 3}});
    HBuf b_r(pixel_locs.data(), halide_type_float, 64, {{npix, 1}, {3, npix}});
    std::vector<double> img((size_t)2 * (size_t)npix);
    HBuf b_img(img.data(), halide_type_float, 64, {{2, 1}, {nu, 2}, {nv, 2 * nu}});
    for (long i = 1; i <= num_iters; ++i) {
        std::cout << "(" << i << ") HIP backprojection start " << std::endl;
        start = steady_clock::now();
        rv = bp(&b_phs, &b_k_r, taylor, N_fft, pd->delta_r, &b_u, &b_v, &b_pos, &b_r, &b_img);
        std::cout << "(" << i << ") HIP backprojection returned " << rv << " in " << ms_since(start)
                  << " ms" << std::endl;
        if (rv != 0) {
            std::cerr << "sarbp: " << sarbp_last_error() << std::endl;
            return rv;
        }
    }
    dft_destroy_fftw();

    std::vector<double> dB((size_t)npix);
    HBuf b_dB(dB.data(), halide_type_float, 64, {{nu, 1}, {nv, nu}});
    std::cout << "HIP dB conversion start" << std::endl;
    start = steady_clock::now();
    rv = img_output_to_dB(&b_img, &b_dB);
    std::cout << "HIP dB conversion returned " << rv << " in " << ms_since(start) << " ms"
              << std::endl;
    if (rv != 0) return rv;

    std::vector<uint8_t> u8((size_t)npix);
    HBuf b_u8(u8.data(), halide_type_uint, 8, {{nu, 1}, {nv, nu}});
    std::cout << "HIP PNG production start" << std::endl;
    start = steady_clock::now();
    rv = img_output_u8(&b_dB, dB_min, dB_max, &b_u8);
    std::cout << "HIP PNG production returned " << rv << " in " << ms_since(start) << " ms"
              << std::endl;
    if (rv != 0) return rv;
    start = steady_clock::now();
    rv = sarbp_png_write_gray8(output_png.c_str(), u8.data(), nu, nv, nu);
    if (rv != 0) {
        std::cerr << "sarbp: " << sarbp_host_last_error() << std::endl;
        return rv;
    }
    std::cout << "Wrote " << output_png << " in " << ms_since(start) << " ms" << std::endl;
    return 0;
}
