// The reference-side binding: a pt::rhi::Device subclass that forwards every
// call to dmt::rhi::HipDevice.  It is written against the reference's own RHI
// headers (src/rhi/Device.h, CommandBuffer.h, Resources.h, Handles.h,
// Types.h, Swapchain.h) and belongs in the reference tree as
// src/rhi_hip/HipDevice.cpp, linked with libdmt_rhi.so, with:
//   - BackendType::Hip added to src/rhi/Types.h:9-14 (kHipBackend below until then),
//   - `case BackendType::Hip: return CreateHipDevice(w);` in Device::Create
//     (src/rhi/Device.cpp:20-48) under PT_HAS_HIP_BACKEND,
//   - "hip" among r_backend's values (src/engine/Engine.cpp:300-302).
// This library does not build it (the headers are the reference's); its test
// compiles it against them.
#include "rhi/CommandBuffer.h"
#include "rhi/Device.h"

#include "dmt_rhi.h"

#include <memory>
#include <vector>

namespace pt::rhi {

// BackendType::Hip once Types.h has it.
constexpr BackendType kHipBackend = static_cast<BackendType>(4);

namespace {

dmt::rhi::BufferHandle D(BufferHandle h) { return {h.id}; }
dmt::rhi::TextureHandle D(TextureHandle h) { return {h.id}; }
dmt::rhi::PipelineHandle D(PipelineHandle h) { return {h.id}; }
dmt::rhi::AccelStructHandle D(AccelStructHandle h) { return {h.id}; }

class HipCommandBufferFwd final : public CommandBuffer {
public:
    explicit HipCommandBufferFwd(dmt::rhi::HipCommandBuffer* cb) : cb_(cb) {}
    dmt::rhi::HipCommandBuffer* native() const { return cb_; }

    void BindComputePipeline(PipelineHandle p) override { cb_->BindComputePipeline(D(p)); }
    void BindBuffer(std::uint32_t slot, BufferHandle b, std::size_t offset) override {
        cb_->BindBuffer(slot, D(b), offset);
    }
    void BindStorageTexture(std::uint32_t slot, TextureHandle t) override { cb_->BindStorageTexture(slot, D(t)); }
    void BindAccelStruct(std::uint32_t slot, AccelStructHandle a) override { cb_->BindAccelStruct(slot, D(a)); }
    void PushConstants(const void* data, std::size_t size) override { cb_->PushConstants(data, size); }
    void Dispatch(std::uint32_t gx, std::uint32_t gy, std::uint32_t gz) override { cb_->Dispatch(gx, gy, gz); }
    void CopyBufferToTexture(BufferHandle src, TextureHandle dst) override {
        cb_->CopyBufferToTexture(D(src), D(dst));
    }
    void ClearStorageTexture(TextureHandle t, const float rgba[4]) override { cb_->ClearStorageTexture(D(t), rgba); }
    void Barrier(const BarrierDesc& d) override {
        dmt::rhi::BarrierDesc b;
        b.from = static_cast<dmt::rhi::BarrierDesc::Stage>(d.from);
        b.to = static_cast<dmt::rhi::BarrierDesc::Stage>(d.to);
        cb_->Barrier(b);
    }

private:
    dmt::rhi::HipCommandBuffer* cb_;
};

class HipDevice final : public Device {
public:
    explicit HipDevice(const NativeWindowHandle& w) {
        dmt::rhi::HipDeviceOptions o;
        // the software device's 1280x720 sentinel for a window not yet sized
        // (SoftwareDevice.cpp:159-164)
        o.width = w.width > 0 ? static_cast<std::uint32_t>(w.width) : 1280u;
        o.height = w.height > 0 ? static_cast<std::uint32_t>(w.height) : 720u;
        dev_ = std::make_unique<dmt::rhi::HipDevice>(o);
        cb_ = std::make_unique<HipCommandBufferFwd>(dev_->AcquireCommandBuffer());
    }

    BufferHandle CreateBuffer(const BufferDesc& d) override {
        dmt::rhi::BufferDesc b;
        b.size = d.size;
        b.usage = static_cast<dmt::rhi::BufferUsage>(d.usage);
        b.debug_name = d.debug_name;
        return BufferHandle{dev_->CreateBuffer(b).id};
    }
    TextureHandle CreateTexture(const TextureDesc& d) override {
        dmt::rhi::TextureDesc t;
        t.width = d.width;
        t.height = d.height;
        t.format = static_cast<dmt::rhi::TextureFormat>(d.format);
        t.usage = static_cast<dmt::rhi::TextureUsage>(d.usage);
        t.debug_name = d.debug_name;
        return TextureHandle{dev_->CreateTexture(t).id};
    }
    PipelineHandle CreateComputePipeline(const ComputePipelineDesc& d) override {
        dmt::rhi::ComputePipelineDesc p;
        p.kernel_name = d.kernel_name;
        p.bytecode = d.bytecode;
        p.debug_name = d.debug_name;
        return PipelineHandle{dev_->CreateComputePipeline(p).id};
    }
    AccelStructHandle CreateBLAS(const BLASDesc& d) override {
        dmt::rhi::BLASDesc b;
        b.vertex_positions = d.vertex_positions;
        b.vertex_count = d.vertex_count;
        b.indices = d.indices;
        b.index_count = d.index_count;
        b.debug_name = d.debug_name;
        return AccelStructHandle{dev_->CreateBLAS(b).id};
    }
    AccelStructHandle CreateTLAS(const TLASDesc& d) override {
        std::vector<dmt::rhi::TLASInstance> v(d.instances.size());
        for (std::size_t i = 0; i < v.size(); ++i) {
            const TLASInstance& s = d.instances[i];
            v[i].blas = D(s.blas);
            for (int k = 0; k < 12; ++k) v[i].transform[k] = s.transform[k];
            v[i].instance_id = s.instance_id;
            v[i].mask = s.mask;
        }
        dmt::rhi::TLASDesc t;
        t.instances = std::span<const dmt::rhi::TLASInstance>(v.data(), v.size());
        t.debug_name = d.debug_name;
        return AccelStructHandle{dev_->CreateTLAS(t).id};
    }
    void DestroyBuffer(BufferHandle h) override { dev_->DestroyBuffer(D(h)); }
    void DestroyTexture(TextureHandle h) override { dev_->DestroyTexture(D(h)); }
    void DestroyPipeline(PipelineHandle h) override { dev_->DestroyPipeline(D(h)); }
    void DestroyAccelStruct(AccelStructHandle h) override { dev_->DestroyAccelStruct(D(h)); }

    void WriteBuffer(BufferHandle h, const void* src, std::size_t size, std::size_t dst_offset) override {
        dev_->WriteBuffer(D(h), src, size, dst_offset);
    }
    bool WriteTexture(TextureHandle h, const void* src, std::size_t src_size) override {
        return dev_->WriteTexture(D(h), src, src_size);
    }

    FrameContext BeginFrame() override {
        const dmt::rhi::FrameContext f = dev_->BeginFrame();
        return FrameContext{TextureHandle{f.swapchain_image.id}, f.width, f.height, f.frame_index};
    }
    void EndFrame(CommandBuffer*) override { dev_->EndFrame(cb_->native()); }
    CommandBuffer* AcquireCommandBuffer() override { return cb_.get(); }
    void Submit(CommandBuffer*) override { dev_->Submit(cb_->native()); }
    void WaitIdle() override { dev_->WaitIdle(); }
    void Resize(int w, int h) override { dev_->Resize(w, h); }

    bool ReadbackTexture(TextureHandle h, void* dst, std::size_t dst_size, std::uint32_t* out_w,
                         std::uint32_t* out_h) override {
        return dev_->ReadbackTexture(D(h), dst, dst_size, out_w, out_h);
    }
    bool ReadbackBuffer(BufferHandle h, void* dst, std::size_t bytes) override {
        return dev_->ReadbackBuffer(D(h), dst, bytes);
    }

    BackendType Type() const override { return kHipBackend; }
    bool SupportsHardwareRT() const override { return dev_->SupportsHardwareRT(); }
    const char* DeviceName() const override { return dev_->DeviceName(); }
    std::size_t CurrentAllocatedBytes() const override { return dev_->CurrentAllocatedBytes(); }
    bool IsDeviceLost() const override { return dev_->IsDeviceLost(); }

private:
    std::unique_ptr<dmt::rhi::HipDevice> dev_;
    std::unique_ptr<HipCommandBufferFwd> cb_;
};

}  // namespace

std::unique_ptr<Device> CreateHipDevice(const NativeWindowHandle& w) { return std::make_unique<HipDevice>(w); }

}  // namespace pt::rhi
