// Python binding of dmt::rhi::HipDevice (include/dmt_rhi.h) for the engine
// call sequences the tests replay: one device object that also carries the
// device's command buffer, numpy arrays in and out.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <vector>

#include "../../include/dmt_rhi.h"

namespace py = pybind11;
using namespace dmt::rhi;

namespace {

std::string as_bytes(const py::buffer& b) {
    py::buffer_info info = b.request();
    const char* p = static_cast<const char*>(info.ptr);
    return std::string(p, p + info.size * info.itemsize);
}

py::dict decoded(const PathtracePush& P) {
    const dmt_frame_params& F = P.frame;
    py::dict d;
    d["pos_fovtan"] = std::vector<float>(F.pos_fovtan, F.pos_fovtan + 4);
    d["fwd_aspect"] = std::vector<float>(F.fwd_aspect, F.fwd_aspect + 4);
    d["right_xyz"] = std::vector<float>(F.right_xyz, F.right_xyz + 4);
    d["up_xyz"] = std::vector<float>(F.up_xyz, F.up_xyz + 4);
    d["frame_index"] = F.frame_index;
    d["spp"] = F.spp;
    d["max_bounces"] = F.max_bounces;
    d["reset_accum"] = F.reset_accum;
    d["mis_enabled"] = F.mis_enabled;
    d["light_tree_enabled"] = F.light_tree_enabled;
    d["refract_bounces"] = F.refract_bounces;
    d["firefly_clamp"] = F.firefly_clamp;
    d["ema_alpha"] = F.ema_alpha;
    d["flags"] = F.flags;
    d["tlas_present"] = P.tlas_present;
    d["prim_count"] = P.prim_count;
    d["light_count"] = P.light_count;
    d["env_map_present"] = P.env_map_present;
    d["env_intensity"] = P.env_intensity;
    d["tonemap_op"] = P.tonemap_op;
    d["motion_blur"] = P.motion_blur;
    d["dof"] = std::vector<float>(F.dof, F.dof + 4);
    d["cat_eye"] = F.cat_eye;
    d["spectral"] = F.spectral;
    d["cauchy_c"] = F.cauchy_c;
    d["mesh_shift_prev"] = std::vector<float>(F.mesh_shift_prev, F.mesh_shift_prev + 3);
    d["mesh_shift_curr"] = std::vector<float>(F.mesh_shift_curr, F.mesh_shift_curr + 3);
    d["mesh_shift_active"] = F.mesh_shift_active;
    d["denoiser"] = P.denoiser;
    d["write_normal"] = P.write_normal;
    d["write_albedo"] = P.write_albedo;
    d["write_specular_albedo"] = P.write_specular_albedo;
    d["write_roughness"] = P.write_roughness;
    d["write_specular_distance"] = P.write_specular_distance;
    d["write_hdr_aux"] = P.write_hdr_aux;
    d["hdr_pipeline"] = P.hdr_pipeline;
    d["clouds_mode"] = P.clouds_mode;
    d["jitter"] = std::vector<float>(F.jitter, F.jitter + 2);
    d["sun"] = std::vector<float>(F.sun, F.sun + 4);
    d["sun_extra"] = std::vector<float>(F.sun_extra, F.sun_extra + 4);
    d["sun_disc_brightness"] = F.sun_disc_brightness;
    d["moon"] = std::vector<float>(F.moon, F.moon + 4);
    d["moon_extra"] = std::vector<float>(F.moon_extra, F.moon_extra + 4);
    d["hdri_light_count"] = F.hdri_light_count;
    std::vector<float> hd, hc;
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 4; ++c) {
            hd.push_back(F.hdri_dir[i][c]);
            hc.push_back(F.hdri_col[i][c]);
        }
    d["hdri_dir"] = hd;
    d["hdri_col"] = hc;
    d["env_total_luminance"] = P.env_total_luminance;
    d["sdf_count"] = P.sdf_count;
    d["sdf_max_iters"] = F.sdf_max_iters;
    d["sdf_epsilon"] = F.sdf_epsilon;
    d["mesh_tex"] = std::vector<uint32_t>(F.mesh_tex, F.mesh_tex + 4);
    d["water_absorb"] = std::vector<float>(F.water_absorb, F.water_absorb + 3);
    d["water_ior"] = F.water_ior;
    d["wave"] = std::vector<float>(F.wave, F.wave + 4);
    d["camera_in_water"] = F.camera_in_water;
    d["deep_color"] = std::vector<float>(F.deep_color, F.deep_color + 3);
    d["curr_view_proj"] = std::vector<float>(P.curr_view_proj, P.curr_view_proj + 16);
    d["prev_view_proj"] = std::vector<float>(P.prev_view_proj, P.prev_view_proj + 16);
    d["ignored"] = P.ignored;
    return d;
}

class PyDevice {
public:
    PyDevice(int device_index, uint32_t width, uint32_t height, bool strict, bool device_bvh_build) {
        HipDeviceOptions o;
        o.device_index = device_index;
        o.width = width;
        o.height = height;
        o.strict = strict;
        o.device_bvh_build = device_bvh_build;
        dev_ = std::make_unique<HipDevice>(o);
        cb_ = dev_->AcquireCommandBuffer();
    }
    HipDevice& d() {
        if (!dev_) throw py::value_error("the device is closed");
        return *dev_;
    }
    HipCommandBuffer& cb() {
        if (!dev_) throw py::value_error("the device is closed");
        return *cb_;
    }
    void close() {
        dev_.reset();
        cb_ = nullptr;
    }

private:
    std::unique_ptr<HipDevice> dev_;
    HipCommandBuffer* cb_ = nullptr;
};

}  // namespace

PYBIND11_MODULE(_dmt_rhi, m) {
    m.doc() = "dmt::rhi::HipDevice, the HIP path tracer behind the engine's RHI calls";
    m.attr("PTPUSH_SIZE") = ptpush::kSize;

    m.def(
        "decode_pathtrace_push",
        [](const py::buffer& push, bool strict) {
            const std::string b = as_bytes(push);
            PathtracePush P;
            std::string err;
            if (!DecodePathtracePush(b.data(), b.size(), strict, &P, &err)) throw py::value_error(err);
            return decoded(P);
        },
        py::arg("push"), py::arg("strict") = true);

    py::class_<PyDevice>(m, "HipDevice")
        .def(py::init<int, uint32_t, uint32_t, bool, bool>(), py::arg("device_index") = 0, py::arg("width") = 1280,
             py::arg("height") = 720, py::arg("strict") = true, py::arg("device_bvh_build") = false)
        .def("close", &PyDevice::close)
        .def("create_buffer", [](PyDevice& s, size_t size) { return s.d().CreateBuffer(BufferDesc{size}).id; })
        .def(
            "write_buffer",
            [](PyDevice& s, uint64_t h, const py::buffer& data, size_t offset) {
                const std::string b = as_bytes(data);
                s.d().WriteBuffer(BufferHandle{h}, b.data(), b.size(), offset);
            },
            py::arg("handle"), py::arg("data"), py::arg("offset") = 0)
        .def(
            "create_texture",
            [](PyDevice& s, uint32_t w, uint32_t h, int format, int usage) {
                TextureDesc t;
                t.width = w;
                t.height = h;
                t.format = static_cast<TextureFormat>(format);
                t.usage = static_cast<TextureUsage>(usage);
                return s.d().CreateTexture(t).id;
            },
            py::arg("width"), py::arg("height"), py::arg("format") = (int)TextureFormat::RGBA32F,
            py::arg("usage") = (int)TextureUsage::Storage)
        .def("write_texture",
             [](PyDevice& s, uint64_t h, const py::buffer& data) {
                 const std::string b = as_bytes(data);
                 return s.d().WriteTexture(TextureHandle{h}, b.data(), b.size());
             })
        .def("create_pipeline",
             [](PyDevice& s, const std::string& name) {
                 ComputePipelineDesc p;
                 p.kernel_name = name;
                 return s.d().CreateComputePipeline(p).id;
             })
        .def("create_blas",
             [](PyDevice& s, py::array_t<float, py::array::c_style | py::array::forcecast> pos,
                py::array_t<uint32_t, py::array::c_style | py::array::forcecast> idx) {
                 BLASDesc b;
                 b.vertex_positions = pos.data();
                 b.vertex_count = (uint32_t)(pos.size() / 3);
                 b.indices = idx.data();
                 b.index_count = (uint32_t)idx.size();
                 return s.d().CreateBLAS(b).id;
             })
        .def("create_tlas",
             [](PyDevice& s, const std::vector<py::tuple>& instances) {
                 // (blas, transform 12 floats or None, instance_id, mask)
                 std::vector<TLASInstance> v;
                 for (const py::tuple& t : instances) {
                     TLASInstance in;
                     in.blas = AccelStructHandle{t[0].cast<uint64_t>()};
                     if (t.size() > 1 && !t[1].is_none()) {
                         auto m = t[1].cast<std::vector<float>>();
                         if (m.size() != 12) throw py::value_error("transform needs 12 floats");
                         std::memcpy(in.transform, m.data(), sizeof(in.transform));
                     }
                     if (t.size() > 2) in.instance_id = t[2].cast<uint32_t>();
                     if (t.size() > 3) in.mask = t[3].cast<uint32_t>();
                     v.push_back(in);
                 }
                 TLASDesc d;
                 d.instances = std::span<const TLASInstance>(v.data(), v.size());
                 return s.d().CreateTLAS(d).id;
             })
        .def("destroy_buffer", [](PyDevice& s, uint64_t h) { s.d().DestroyBuffer(BufferHandle{h}); })
        .def("destroy_texture", [](PyDevice& s, uint64_t h) { s.d().DestroyTexture(TextureHandle{h}); })
        .def("destroy_pipeline", [](PyDevice& s, uint64_t h) { s.d().DestroyPipeline(PipelineHandle{h}); })
        .def("destroy_accel", [](PyDevice& s, uint64_t h) { s.d().DestroyAccelStruct(AccelStructHandle{h}); })
        .def("begin_frame",
             [](PyDevice& s) {
                 FrameContext f = s.d().BeginFrame();
                 return py::make_tuple(f.swapchain_image.id, f.width, f.height, f.frame_index);
             })
        .def("end_frame", [](PyDevice& s) { s.d().EndFrame(&s.cb()); })
        .def("submit", [](PyDevice& s) { s.d().Submit(&s.cb()); })
        .def("wait_idle", [](PyDevice& s) { s.d().WaitIdle(); })
        .def("resize", [](PyDevice& s, int w, int h) { s.d().Resize(w, h); })
        .def("bind_pipeline", [](PyDevice& s, uint64_t h) { s.cb().BindComputePipeline(PipelineHandle{h}); })
        .def(
            "bind_buffer",
            [](PyDevice& s, uint32_t slot, uint64_t h, size_t off) { s.cb().BindBuffer(slot, BufferHandle{h}, off); },
            py::arg("slot"), py::arg("handle"), py::arg("offset") = 0)
        .def("bind_texture",
             [](PyDevice& s, uint32_t slot, uint64_t h) { s.cb().BindStorageTexture(slot, TextureHandle{h}); })
        .def("bind_accel",
             [](PyDevice& s, uint32_t slot, uint64_t h) { s.cb().BindAccelStruct(slot, AccelStructHandle{h}); })
        .def("push_constants",
             [](PyDevice& s, const py::buffer& data) {
                 const std::string b = as_bytes(data);
                 s.cb().PushConstants(b.data(), b.size());
             })
        .def(
            "dispatch", [](PyDevice& s, uint32_t gx, uint32_t gy, uint32_t gz) { s.cb().Dispatch(gx, gy, gz); },
            py::arg("gx") = 1, py::arg("gy") = 1, py::arg("gz") = 1)
        .def("clear_texture",
             [](PyDevice& s, uint64_t h, std::vector<float> rgba) {
                 if (rgba.size() != 4) throw py::value_error("rgba needs 4 floats");
                 s.cb().ClearStorageTexture(TextureHandle{h}, rgba.data());
             })
        .def("copy_buffer_to_texture",
             [](PyDevice& s, uint64_t b, uint64_t t) { s.cb().CopyBufferToTexture(BufferHandle{b}, TextureHandle{t}); })
        .def("readback_texture",
             [](PyDevice& s, uint64_t h) -> py::object {
                 uint32_t w = 0, hh = 0;
                 // size query: a zero-byte destination fails after reporting the size
                 char probe = 0;
                 (void)s.d().ReadbackTexture(TextureHandle{h}, &probe, 0, &w, &hh);
                 py::array_t<float> out({(py::ssize_t)hh, (py::ssize_t)w, (py::ssize_t)4});
                 if (!s.d().ReadbackTexture(TextureHandle{h}, out.mutable_data(), (size_t)w * hh * 16u, &w, &hh))
                     return py::none();
                 return out;
             })
        .def("readback_buffer",
             [](PyDevice& s, uint64_t h, size_t n) -> py::object {
                 std::string b(n, '\0');
                 if (!s.d().ReadbackBuffer(BufferHandle{h}, b.data(), n)) return py::none();
                 return py::bytes(b);
             })
        .def("is_device_lost", [](PyDevice& s) { return s.d().IsDeviceLost(); })
        .def("supports_hardware_rt", [](PyDevice& s) { return s.d().SupportsHardwareRT(); })
        .def("device_name", [](PyDevice& s) { return std::string(s.d().DeviceName()); })
        .def("allocated_bytes", [](PyDevice& s) { return s.d().CurrentAllocatedBytes(); })
        .def("last_error", [](PyDevice& s) { return s.d().LastError(); })
        .def("ignored_features", [](PyDevice& s) { return s.d().IgnoredFeatures(); })
        .def("scene_builds", [](PyDevice& s) { return s.d().SceneBuilds(); })
        .def("material_updates", [](PyDevice& s) { return s.d().MaterialUpdates(); })
        .def("mesh_refits", [](PyDevice& s) { return s.d().MeshRefits(); })
        .def("active_scene", [](PyDevice& s) { return reinterpret_cast<uintptr_t>(s.d().ActiveScene()); });
}
