// SHA-256 (FIPS 180-4), used for checkpoint checksums and scene identities.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>

namespace rp {

class Sha256 {
  public:
    Sha256();
    void update(const void *data, size_t n);
    std::array<uint8_t, 32> finish();

  private:
    void block(const uint8_t *p);
    uint32_t h_[8];
    uint8_t buf_[64];
    size_t fill_ = 0;
    uint64_t bytes_ = 0;
};

} // namespace rp
