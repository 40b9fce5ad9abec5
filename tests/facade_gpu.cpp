// Built by build() into tests/_bin/facade_gpu and run by
// tests/test_gpu_facade.py: the C++ facade (orb_gpu.hpp) on the device, the
// way Frame's stereo constructor uses ORB-SLAM2's classes (Frame.cc:146-169):
// two extractors, operator() on the left and right images, then
// ComputeStereoMatches, UndistortKeyPoints and ComputeStereoFromRGBD; also
// the pyramid getter, the scale getters, the
// empty-image convention and the error channel.
//
// usage: facade_gpu left.raw right.raw width height nfeatures bf b outdir
// writes outdir/{kl,dl,kr,dr,ur,depth,pyr,scales}.bin and prints "facade gpu ok".
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "orb_gpu.hpp"

static std::vector<uint8_t> read_file(const std::string& path, size_t n)
{
    std::vector<uint8_t> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), 1, n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

static void write_file(const std::string& path, const void* p, size_t n)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || (n && std::fwrite(p, 1, n, f) != n)) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(3);
    }
    std::fclose(f);
}

int main(int argc, char** argv)
{
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s left.raw right.raw width height nfeatures bf b outdir\n", argv[0]);
        return 1;
    }
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), nf = std::atoi(argv[5]);
    const float bf = std::strtof(argv[6], nullptr), b = std::strtof(argv[7], nullptr);
    const std::string out = argv[8];
    const std::vector<uint8_t> imL = read_file(argv[1], (size_t)w * h), imR = read_file(argv[2], (size_t)w * h);

    orbgpu::Context ctx(0);
    orbgpu::ORBextractor left(ctx, nf, 1.2f, 8, 20, 7), right(ctx, nf, 1.2f, 8, 20, 7);

    std::vector<orb_keypoint> kl, kr;
    std::vector<uint8_t> dl, dr;
    left({imL.data(), w, h, (size_t)w}, nullptr, kl, dl);
    right({imR.data(), w, h, (size_t)w}, nullptr, kr, dr);
    std::vector<float> ur, depth;
    orbgpu::ComputeStereoMatches(left, right, kl, dl, kr, dr, bf, b, ur, depth);

    // The RGB-D constructor's steps on the left keypoints: TUM fr1's model,
    // and a depth ramp with a hole (the test restates both).
    const float K[4] = {517.306408f, 516.469215f, 318.643040f, 255.313989f};
    const std::vector<float> D = {0.262383f, -0.953104f, -0.005358f, 0.002628f, 1.163314f};
    std::vector<orb_keypoint> kun;
    orbgpu::UndistortKeyPoints(ctx, kl, K, D, kun);
    std::vector<float> dimg((size_t)w * h);
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++)
            dimg[(size_t)r * w + c] = (r >= 100 && r < 160 && c >= 300 && c < 380) ? 0.0f : 1.0f + (float)(r + c) / 400.0f;
    std::vector<float> rur, rdep;
    orbgpu::ComputeStereoFromRGBD(ctx, kl, kun, dimg.data(), w, h, (size_t)w * 4, bf, rur, rdep);

    // mvImagePyramid of the left extractor, levels concatenated
    std::vector<uint8_t> pyr;
    for (int l = 0; l < left.GetLevels(); l++) {
        int lw = 0, lh = 0;
        const std::vector<uint8_t> px = left.ImagePyramidLevel(l, &lw, &lh);
        pyr.insert(pyr.end(), px.begin(), px.end());
    }
    std::vector<float> scales = left.GetScaleFactors();
    for (const std::vector<float>& t : {left.GetInverseScaleFactors(), left.GetScaleSigmaSquares(),
                                        left.GetInverseScaleSigmaSquares()})
        scales.insert(scales.end(), t.begin(), t.end());

    write_file(out + "/kl.bin", kl.data(), kl.size() * sizeof(orb_keypoint));
    write_file(out + "/dl.bin", dl.data(), dl.size());
    write_file(out + "/kr.bin", kr.data(), kr.size() * sizeof(orb_keypoint));
    write_file(out + "/dr.bin", dr.data(), dr.size());
    write_file(out + "/ur.bin", ur.data(), ur.size() * sizeof(float));
    write_file(out + "/depth.bin", depth.data(), depth.size() * sizeof(float));
    write_file(out + "/pyr.bin", pyr.data(), pyr.size());
    write_file(out + "/scales.bin", scales.data(), scales.size() * sizeof(float));
    write_file(out + "/kun.bin", kun.data(), kun.size() * sizeof(orb_keypoint));
    write_file(out + "/rgbd_ur.bin", rur.data(), rur.size() * sizeof(float));
    write_file(out + "/rgbd_depth.bin", rdep.data(), rdep.size() * sizeof(float));

    // An empty image returns with the outputs untouched (ORBextractor.cc:1537).
    std::vector<orb_keypoint> k0 = kl;
    std::vector<uint8_t> d0 = dl;
    left({nullptr, 0, 0, 0}, nullptr, k0, d0);
    if (k0.size() != kl.size() || d0 != dl) return 4;
    // Errors surface as orbgpu::Error with the C ABI's code.
    try {
        orbgpu::ORBextractor bad(ctx, 1000, 1.0f, 8, 20, 7);   // scaleFactor must exceed 1
        return 5;
    } catch (const orbgpu::Error& e) {
        if (e.code != ORB_E_ARG) return 6;
    }
    std::printf("facade gpu ok: %zu + %zu keypoints\n", kl.size(), kr.size());
    return 0;
}
