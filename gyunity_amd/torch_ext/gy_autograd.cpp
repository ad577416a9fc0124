// LibTorch binding of the differentiable render: torch::autograd::Functions
// whose forward renders the scene with the texture texels given as a tensor and
// whose backward runs the path-replay adjoint (gy_adjoint_device for one
// view, gy_batch_adjoint_device for the views of a batch in one launch).
//
//   image = RenderTextures::apply(textures, scene, camera, params)
//     textures: float32 CUDA tensor, one value per texture float of the scene
//               (the layout of gy_scene_desc.tex_data)
//     image:    float32 CUDA tensor (height, width, 3), colour / weight
//
// The library runs on its own stream and synchronises at the end of every
// call; the device is synchronised before each call so that torch's pending
// work on the inputs is complete.
#include <torch/extension.h>
#include <torch/cuda.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include <gyunity.h>

namespace {

void Check(int rc, const char* what) {
    if (rc != GY_OK) throw std::runtime_error(std::string(what) + ": " + gy_last_error());
}

gy_camera_desc CameraOf(const std::vector<double>& c) {
    // pos(3), look_at(3), up(3), fovy, film_dist, width, height
    if (c.size() != 13) throw std::invalid_argument("camera: 13 values expected");
    gy_camera_desc d;
    for (int k = 0; k < 3; ++k) {
        d.pos[k] = (float)c[k];
        d.look_at[k] = (float)c[3 + k];
        d.up[k] = (float)c[6 + k];
    }
    d.fovy = (float)c[9];
    d.film_dist = (float)c[10];
    d.width = (int32_t)c[11];
    d.height = (int32_t)c[12];
    return d;
}

gy_render_params ParamsOf(const std::vector<int64_t>& p) {
    // spp, max_depth, rr_depth, mode, samples_per_pass, block_size, refill_threshold, sampler, ray_sort
    if (p.size() != 9) throw std::invalid_argument("params: 9 values expected");
    gy_render_params r = {};
    r.spp = (int32_t)p[0];
    r.max_depth = (int32_t)p[1];
    r.rr_depth = (int32_t)p[2];
    r.mode = (int32_t)p[3];
    r.samples_per_pass = (int32_t)p[4];
    r.block_size = (int32_t)p[5];
    r.refill_threshold = (int32_t)p[6];
    r.sampler = (int32_t)p[7];
    r.ray_sort = (int32_t)p[8];
    return r;
}

struct RenderTextures : public torch::autograd::Function<RenderTextures> {
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor textures, int64_t scene,
                                 std::vector<double> camera, std::vector<int64_t> params) {
        gy_scene* sc = reinterpret_cast<gy_scene*>(scene);
        TORCH_CHECK(textures.is_cuda() && textures.scalar_type() == torch::kFloat32, "textures: float32 CUDA tensor");
        TORCH_CHECK(textures.numel() == gy_scene_texture_floats(sc), "textures: one value per texture float of the scene");
        const gy_camera_desc cam = CameraOf(camera);
        const gy_render_params prm = ParamsOf(params);
        torch::Tensor tex = textures.contiguous();
        torch::cuda::synchronize();
        Check(gy_scene_set_textures_device(sc, tex.data_ptr<float>(), tex.numel()), "gy_scene_set_textures_device");
        Check(gy_render(sc, &cam, &prm, nullptr, nullptr), "gy_render");
        torch::Tensor image = torch::empty({cam.height, cam.width, 3}, textures.options());
        Check(gy_film_image_device(sc, image.data_ptr<float>()), "gy_film_image_device");
        ctx->saved_data["scene"] = scene;
        ctx->saved_data["camera"] = camera;
        ctx->saved_data["params"] = params;
        ctx->saved_data["n"] = textures.numel();
        return image;
    }

    static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx, torch::autograd::tensor_list grads) {
        gy_scene* sc = reinterpret_cast<gy_scene*>(ctx->saved_data["scene"].toInt());
        const gy_camera_desc cam = CameraOf(ctx->saved_data["camera"].toDoubleVector());
        const gy_render_params prm = ParamsOf(ctx->saved_data["params"].toIntVector());
        torch::Tensor dimage = grads[0].to(torch::kFloat32).contiguous();
        torch::Tensor dtex = torch::zeros({ctx->saved_data["n"].toInt()}, dimage.options());
        torch::cuda::synchronize();
        // the film still holds this forward's render unless another render of
        // the scene ran in between; gy_adjoint_device refuses a mismatch
        Check(gy_adjoint_device(sc, &cam, &prm, dimage.data_ptr<float>(), dtex.data_ptr<float>()), "gy_adjoint_device");
        return {dtex, torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

// Many views of one textured scene as one batch: forward renders every entry
// of the batch in one lane space (gy_batch_render) and returns the entries'
// images concatenated (gy_batch_images_device); backward replays every sample
// of every entry in one launch (gy_batch_adjoint_device).
struct RenderViews : public torch::autograd::Function<RenderViews> {
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor textures, int64_t batch, int64_t scene) {
        gy_scene* sc = reinterpret_cast<gy_scene*>(scene);
        gy_batch* b = reinterpret_cast<gy_batch*>(batch);
        TORCH_CHECK(textures.is_cuda() && textures.scalar_type() == torch::kFloat32, "textures: float32 CUDA tensor");
        TORCH_CHECK(textures.numel() == gy_scene_texture_floats(sc), "textures: one value per texture float of the scene");
        torch::Tensor tex = textures.contiguous();
        torch::cuda::synchronize();
        Check(gy_scene_set_textures_device(sc, tex.data_ptr<float>(), tex.numel()), "gy_scene_set_textures_device");
        Check(gy_batch_render(b), "gy_batch_render");
        const int64_t n = gy_batch_image_floats(b);
        TORCH_CHECK(n > 0, "batch without images");
        torch::Tensor images = torch::empty({n}, textures.options());
        Check(gy_batch_images_device(b, images.data_ptr<float>()), "gy_batch_images_device");
        ctx->saved_data["batch"] = batch;
        ctx->saved_data["n"] = textures.numel();
        return images;
    }

    static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx, torch::autograd::tensor_list grads) {
        gy_batch* b = reinterpret_cast<gy_batch*>(ctx->saved_data["batch"].toInt());
        torch::Tensor dimages = grads[0].to(torch::kFloat32).contiguous();
        torch::Tensor dtex = torch::zeros({ctx->saved_data["n"].toInt()}, dimages.options());
        torch::cuda::synchronize();
        // the batch films hold this forward's render unless the batch or the
        // textures changed in between; gy_batch_adjoint_device refuses that
        Check(gy_batch_adjoint_device(b, dimages.data_ptr<float>(), dtex.data_ptr<float>(), nullptr), "gy_batch_adjoint_device");
        return {dtex, torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor render_textures(torch::Tensor textures, int64_t scene, std::vector<double> camera, std::vector<int64_t> params) {
    return RenderTextures::apply(textures, scene, camera, params);
}

torch::Tensor render_views(torch::Tensor textures, int64_t batch, int64_t scene) {
    return RenderViews::apply(textures, batch, scene);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "Differentiable path-traced render of texture texels (path-replay adjoint)";
    m.def("render_textures", &render_textures, "image = render(textures); backward by path replay",
          py::arg("textures"), py::arg("scene"), py::arg("camera"), py::arg("params"));
    m.def("render_views", &render_views, "images = render every view of the batch(textures); backward by path replay "
          "of every view in one launch", py::arg("textures"), py::arg("batch"), py::arg("scene"));
}
