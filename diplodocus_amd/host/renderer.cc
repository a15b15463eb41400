// GpuRenderer and Framebuffer of the host interface (renderer.h): the reference's
// src/renderer/gpu_renderer.cc:14-64 flow (resize the framebuffer, set up the scene on the
// device, launch, download) over the engine's GPU BVH and Whitted kernel.
#include "renderer.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <stdexcept>

namespace diplodocus_amd {

Framebuffer::Framebuffer(int width, int height) { Resize(width, height); }

Vec3 Framebuffer::GetPixel(int x, int y) const {
    if (x < 0 || x >= width_ || y < 0 || y >= height_) throw std::out_of_range("Framebuffer: pixel out of bounds");
    const size_t i = (size_t(y) * size_t(width_) + size_t(x)) * 3;
    return {data_[i], data_[i + 1], data_[i + 2]};
}

void Framebuffer::Clear(const Vec3& color) {
    for (size_t i = 0; i < data_.size(); i += 3) data_[i] = color.x, data_[i + 1] = color.y, data_[i + 2] = color.z;
}

void Framebuffer::Resize(int width, int height) {
    if (width < 0 || height < 0) throw std::invalid_argument("Framebuffer: negative size");
    width_ = width;
    height_ = height;
    data_.resize(size_t(width) * size_t(height) * 3);
}

void Framebuffer::SetPixel(int x, int y, const Vec3& c) {
    if (x < 0 || x >= width_ || y < 0 || y >= height_) throw std::out_of_range("Framebuffer: pixel out of bounds");
    const size_t i = (size_t(y) * size_t(width_) + size_t(x)) * 3;
    data_[i] = c.x, data_[i + 1] = c.y, data_[i + 2] = c.z;
}

RenderResult GpuRenderer::StartRender(const RenderConfig& render_config,
                                      const AccelerationStructureConfig& acceleration_config, const Scene& scene,
                                      Framebuffer& framebuffer, Stats& stats) {
    progress_ = 0.0f;
    last_error_.clear();
    try {
        if (render_config.width < 1 || render_config.height < 1)
            throw std::invalid_argument("render size must be at least 1 x 1");
        if (scene.Triangles().empty()) throw std::invalid_argument("the scene has no triangles");
        framebuffer.Resize(render_config.width, render_config.height);
        framebuffer.Clear(render_config.background_color);
        if (cancelled_) return RenderResult::kCanceled;
        if (!ctx_) ctx_ = std::make_shared<Context>(0);

        AccelerationStructureConfig ac = acceleration_config;
        switch (ac.acceleration_structure_type) {
            case AccelerationStructureType::kPloc:
            case AccelerationStructureType::kPlocEmc:
            case AccelerationStructureType::kPlocSobb:
            case AccelerationStructureType::kPlocEmcSobb:
                break;
            default:  // the GPU builds PLOC trees only; the image does not depend on the tree
                ac.acceleration_structure_type = AccelerationStructureType::kPloc;
                break;
        }
        const std::unique_ptr<AccelerationStructure> as = CreateGpuAccelerationStructure(ac, ctx_);
        as->Build(ac, stats, scene.Triangles());
        if (cancelled_) return RenderResult::kCanceled;

        const GpuBvh& bvh = static_cast<const GpuBvh&>(*as);
        const std::vector<float> img = GpuWhittedRenderer(bvh).Render(render_config, scene.GetCamera(),
                                                                      scene.Triangles(), scene.Materials(), &stats);
        std::memcpy(framebuffer.GetDataPtr(), img.data(), img.size() * sizeof(float));
        progress_ = 1.0f;
        return RenderResult::kDone;
    } catch (const std::exception& e) {
        last_error_ = e.what();
        return RenderResult::kError;
    }
}

}  // namespace diplodocus_amd
