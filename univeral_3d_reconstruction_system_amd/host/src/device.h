// Process-wide device context shared by the drop-in classes, and host<->device helpers.
#pragma once

#include "u3d_hip.h"
#include "u3d/splat/GaussianModel.h"

#include <cstddef>
#include <string>

namespace u3d {
namespace dev {

// The context of the device selected by U3D_DEVICE (default 0); aborts with a message when no
// usable GPU is present -- there is no CPU fallback.
u3d_ctx *context();
[[noreturn]] void fail(const char *what, int rc);
inline void check(int rc, const char *what)
{
    if (rc != U3D_OK)
        fail(what, rc);
}

// A growable device copy of a GaussianModel in SoA layout (capacity >= size).
class Model
{
public:
    ~Model();
    void upload(const GaussianModel &m);          // resizes to m.size()
    void zero(std::size_t n);                      // n zeroed entries
    void download(GaussianModel &m) const;         // m resized to size()
    void downloadAdd(GaussianModel &m) const;      // m += device values
    u3d_gaussians view() const;
    std::size_t size() const { return m_n; }

private:
    void reserve(std::size_t n);
    float *m_buf = nullptr;
    std::size_t m_cap = 0, m_n = 0;
};

// Growable device buffer of bytes.
class Buffer
{
public:
    ~Buffer();
    void *ensure(std::size_t bytes);
    void *get() const { return m_p; }

private:
    void *m_p = nullptr;
    std::size_t m_cap = 0;
};

} // namespace dev
} // namespace u3d
