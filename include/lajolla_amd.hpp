/*
 * lajolla_amd.hpp -- C++ host API over the C ABI, with the shape of lajolla's
 * own entry points:
 *
 *   std::unique_ptr<Scene> parse_scene(path)     src/parsers/parse_scene.h:9
 *   Image3 render(const Scene &)                 src/render.h:9
 *   void imwrite(path, const Image3 &)           src/image.h:57
 *
 * Header-only; link with -llajolla_amd.  Failures throw lajolla_amd::Error
 * (as the reference's Error() throws fl_exception, src/flexception.h:24).
 */
#ifndef LAJOLLA_AMD_HPP
#define LAJOLLA_AMD_HPP

#include "lajolla_amd.h"

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lajolla_amd {

class Error : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

inline void check(int rc) {
    if (rc != LJX_OK) {
        char buf[1024];
        ljx_last_error(buf, sizeof(buf));
        throw Error(buf);
    }
}

/* RGB image, row-major, three floats per pixel (src/image.h:10-40). */
struct Image3 {
    int width = 0, height = 0;
    std::vector<float> data;
    Image3() = default;
    Image3(int w, int h) : width(w), height(h), data(size_t(w) * h * 3, 0.f) {}
    float *operator()(int x, int y) { return &data[(size_t(y) * width + x) * 3]; }
    const float *operator()(int x, int y) const { return &data[(size_t(y) * width + x) * 3]; }
};

/* A scene resident on one device (src/scene.h:42-82).  Non-copyable. */
class Scene {
  public:
    Scene(void *handle, int width, int height, std::string output_filename)
        : h_(handle), width_(width), height_(height), output_filename(std::move(output_filename)) {}
    ~Scene() {
        if (h_) ljx_scene_destroy(h_);
    }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    void *handle() const { return h_; }
    int width() const { return width_; }
    int height() const { return height_; }

  private:
    void *h_;
    int width_, height_;

  public:
    std::string output_filename;  /* the film's filename (src/scene.h:74) */
};

inline std::unique_ptr<Scene> parse_scene(const std::string &filename, int device = 0) {
    ljx_scene_desc *d = nullptr;
    check(ljx_scene_parse(filename.c_str(), &d));
    void *h = nullptr;
    int rc = ljx_scene_create(d, device, &h);
    int w = d->camera.width, hgt = d->camera.height;
    ljx_scene_desc_free(d);
    check(rc);
    std::string out = "image.pfm";
    return std::make_unique<Scene>(h, w, hgt, out);
}

inline Image3 render(const Scene &scene) {
    Image3 img(scene.width(), scene.height());
    check(ljx_render(scene.handle(), nullptr, nullptr, img.data.data(), nullptr));
    return img;
}

/* Portable float map (.pfm): little-endian, rows bottom to top. */
inline void imwrite(const std::string &filename, const Image3 &img) {
    FILE *f = std::fopen(filename.c_str(), "wb");
    if (!f) throw Error("cannot write " + filename);
    std::fprintf(f, "PF\n%d %d\n-1\n", img.width, img.height);
    for (int y = img.height - 1; y >= 0; y--)
        std::fwrite(img(0, y), sizeof(float), size_t(img.width) * 3, f);
    std::fclose(f);
}

}  // namespace lajolla_amd

#endif
