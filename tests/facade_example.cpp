// Compiled by tests/test_facade.py: the C++ facade builds against the C ABI
// and its device-free calls behave as the reference's.
#include <cstdio>
#include <cstring>

#include "orb_gpu.hpp"

int main()
{
    uint8_t a[32], b[32];
    for (int i = 0; i < 32; i++) {
        a[i] = (uint8_t)(i * 37);
        b[i] = (uint8_t)~a[i];
    }
    if (orbgpu::ORBmatcher::DescriptorDistance(a, a) != 0) return 1;
    if (orbgpu::ORBmatcher::DescriptorDistance(a, b) != 256) return 2;
    b[3] ^= 0x81;
    if (orbgpu::ORBmatcher::DescriptorDistance(a, b) != 254) return 3;
    if (orb_extractor_create(nullptr, 1000, 1.2f, 8, 20, 7, nullptr) != ORB_E_ARG) return 4;
    static_assert(sizeof(orb_keypoint) == 28, "cv::KeyPoint layout");
    std::printf("facade ok\n");
    return 0;
}
