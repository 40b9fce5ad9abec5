// orb_gpu.hpp -- C++ facade over the C ABI (orb_gpu.h) with the shapes of
// ORB-SLAM2's own classes, so Tracking/Frame code reads the same:
//
//   ORBextractor(nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST)
//       include/ORBextractor.h:93
//   void operator()(image, mask, keypoints, descriptors)
//       include/ORBextractor.h:111 -- image is a GrayImage view (CV_8UC1
//       pixels, width, height, row stride); the mask is ignored, as in the
//       reference; descriptors are N x 32 bytes, row i for keypoints[i]
//   GetLevels / GetScaleFactor / GetScaleFactors / GetInverseScaleFactors /
//   GetScaleSigmaSquares / GetInverseScaleSigmaSquares
//       include/ORBextractor.h:119-159
//   ImagePyramidLevel(level)   mvImagePyramid[level], include/ORBextractor.h:162
//   ORBmatcher::DescriptorDistance(a, b)
//       include/ORBmatcher.h:65
//   ComputeStereoMatches(left, right, ...)   Frame::ComputeStereoMatches,
//       include/Frame.h:249
//   UndistortKeyPoints(ctx, mvKeys, K, mDistCoef, mvKeysUn)
//       Frame::UndistortKeyPoints, src/Frame.cc:724-776
//   ComputeStereoFromRGBD(ctx, mvKeys, mvKeysUn, depth, mbf, ...)
//       Frame::ComputeStereoFromRGBD, src/Frame.cc:1085-1123
//
// With OpenCV present, cv::Mat / cv::KeyPoint convert directly: orb_keypoint
// has cv::KeyPoint's layout, and a CV_8UC1 Mat is (data, cols, rows, step).
// Errors surface as orbgpu::Error (the reference has no error channel; a bad
// argument there is undefined behaviour).
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "orb_gpu.h"

namespace orbgpu {

struct Error : std::runtime_error {
    int code;
    Error(const char* what, int rc) : std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + ")"),
                                      code(rc) {}
};

inline void check(const char* what, int rc)
{
    if (rc != ORB_OK) throw Error(what, rc);
}

struct GrayImage {
    const uint8_t* data;
    int width, height;
    size_t stride;
};

class Context {
public:
    explicit Context(int device = 0) { check("orb_context_create", orb_context_create(device, &h_)); }
    ~Context() { orb_context_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    orb_context* get() const { return h_; }

private:
    orb_context* h_ = nullptr;
};

class ORBextractor {
public:
    ORBextractor(Context& ctx, int nfeatures, float scaleFactor, int nlevels, int iniThFAST, int minThFAST)
        : nlevels_(nlevels)
    {
        check("orb_extractor_create",
              orb_extractor_create(ctx.get(), nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST, &h_));
    }
    ~ORBextractor() { orb_extractor_destroy(h_); }
    ORBextractor(const ORBextractor&) = delete;
    ORBextractor& operator=(const ORBextractor&) = delete;

    // ORBextractor::operator()(image, mask, keypoints, descriptors)
    void operator()(const GrayImage& image, const void* /*mask: ignored*/, std::vector<orb_keypoint>& keypoints,
                    std::vector<uint8_t>& descriptors)
    {
        if (image.width == 0 || image.height == 0) return;   // outputs untouched, ORBextractor.cc:1537
        keypoints.clear();
        descriptors.clear();
        const int cap = orb_extractor_capacity(h_, image.width, image.height);
        check("orb_extractor_capacity", cap < 0 ? cap : ORB_OK);
        keypoints.resize(cap);
        descriptors.resize((size_t)cap * 32);
        int n = 0;
        check("orb_extract", orb_extract(h_, image.data, image.width, image.height, image.stride, keypoints.data(),
                                         cap, descriptors.data(), &n));
        keypoints.resize(n);
        descriptors.resize((size_t)n * 32);
    }

    int GetLevels() const { return orb_extractor_levels(h_); }
    float GetScaleFactor() const { return orb_extractor_scale_factor(h_); }
    std::vector<float> GetScaleFactors() const { return table(0); }
    std::vector<float> GetInverseScaleFactors() const { return table(1); }
    std::vector<float> GetScaleSigmaSquares() const { return table(2); }
    std::vector<float> GetInverseScaleSigmaSquares() const { return table(3); }

    // mvImagePyramid[level] of the last call, copied to host memory.
    std::vector<uint8_t> ImagePyramidLevel(int level, int* width, int* height)
    {
        check("orb_extractor_pyramid_level", orb_extractor_pyramid_level(h_, level, nullptr, 0, width, height));
        std::vector<uint8_t> px((size_t)*width * *height);
        check("orb_extractor_pyramid_level",
              orb_extractor_pyramid_level(h_, level, px.data(), (size_t)*width, nullptr, nullptr));
        return px;
    }

    orb_extractor* get() const { return h_; }

private:
    std::vector<float> table(int which) const
    {
        std::vector<float> t(nlevels_);
        float* p[4] = {nullptr, nullptr, nullptr, nullptr};
        p[which] = t.data();
        check("orb_extractor_scales", orb_extractor_scales(h_, p[0], p[1], p[2], p[3]));
        return t;
    }
    orb_extractor* h_ = nullptr;
    int nlevels_;
};

struct ORBmatcher {
    static const int TH_HIGH = 100;
    static const int TH_LOW = 50;
    static const int HISTO_LENGTH = 30;
    // ORBmatcher::DescriptorDistance(a, b) on two 32-byte descriptor rows.
    static int DescriptorDistance(const uint8_t* a, const uint8_t* b) { return orb_descriptor_distance(a, b); }
};

// Frame::ComputeStereoMatches for the keypoints the two extractors returned
// on their last calls.  mvuRight / mvDepth get one entry per left keypoint.
inline void ComputeStereoMatches(ORBextractor& left, ORBextractor& right, const std::vector<orb_keypoint>& kl,
                                 const std::vector<uint8_t>& dl, const std::vector<orb_keypoint>& kr,
                                 const std::vector<uint8_t>& dr, float bf, float b, std::vector<float>& mvuRight,
                                 std::vector<float>& mvDepth)
{
    mvuRight.assign(kl.size(), -1.0f);
    mvDepth.assign(kl.size(), -1.0f);
    check("orb_stereo_match", orb_stereo_match(left.get(), right.get(), kl.data(), dl.data(), (int)kl.size(),
                                               kr.data(), dr.data(), (int)kr.size(), bf, b, mvuRight.data(),
                                               mvDepth.data()));
}

// Frame::UndistortKeyPoints: K = fx, fy, cx, cy; D = mDistCoef's 4 or 5
// coefficients (k1 k2 p1 p2 [k3]); k1 == 0 copies.
inline void UndistortKeyPoints(Context& ctx, const std::vector<orb_keypoint>& mvKeys, const float K[4],
                               const std::vector<float>& D, std::vector<orb_keypoint>& mvKeysUn)
{
    mvKeysUn.resize(mvKeys.size());
    check("orb_undistort_keypoints", orb_undistort_keypoints(ctx.get(), (int)mvKeys.size(), mvKeys.data(), K,
                                                             D.data(), (int)D.size(), mvKeysUn.data()));
}

// Frame::ComputeStereoFromRGBD on a float depth image (CV_32F, as Tracking
// scales it): mvuRight / mvDepth get one entry per keypoint (-1 without depth).
inline void ComputeStereoFromRGBD(Context& ctx, const std::vector<orb_keypoint>& mvKeys,
                                  const std::vector<orb_keypoint>& mvKeysUn, const float* depth, int width,
                                  int height, size_t stride, float mbf, std::vector<float>& mvuRight,
                                  std::vector<float>& mvDepth)
{
    mvuRight.assign(mvKeys.size(), -1.0f);
    mvDepth.assign(mvKeys.size(), -1.0f);
    check("orb_rgbd_depth", orb_rgbd_depth(ctx.get(), (int)mvKeys.size(), mvKeys.data(), mvKeysUn.data(), depth, width,
                                           height, stride, mbf, mvuRight.data(), mvDepth.data()));
}

}  // namespace orbgpu
