// Renders the Cornell box through gy::HipPathTracing (the Integrator
// interface) and checks the film against direct C-ABI renders: the adapter
// must add nothing, and the megakernel and wavefront films must be equal bit
// for bit.  Exit status 0 on success.
//
//   render_cornell [--expect-no-device]
// With --expect-no-device the program checks that opening a device fails with
// a gy::Error (hosts without a GPU) and exits 0.
#include <gyunity/HipPathTracing.h>

#include <cmath>
#include <cstdio>
#include <cstring>

int main(int argc, char** argv) {
    const bool expect_no_device = argc > 1 && std::strcmp(argv[1], "--expect-no-device") == 0;
    std::shared_ptr<gy::Device> dev;
    try {
        dev = std::make_shared<gy::Device>(0);
    } catch (const gy::Error& e) {
        std::printf("no device: %s\n", e.what());
        return expect_no_device ? 0 : 1;
    }
    if (expect_no_device) {
        std::printf("a device opened although none was expected\n");
        return 1;
    }
    try {
        gy_generated* g = nullptr;
        gy::Check(gy_generate_scene(GY_SCENE_CORNELL, 0, 0, &g), "gy_generate_scene");
        gy_scene_desc desc;
        gy_camera_desc cd;
        gy_render_params rp;
        gy::Check(gy_generated_desc(g, &desc, &cd, &rp), "gy_generated_desc");
        const gy::Scene scene = gy::Scene::FromDesc(desc);
        gy_generated_free(g);

        const int W = 128, H = 128;
        auto film = std::make_shared<gy::Film>(W, H);
        gy::PinHoleCamera camera(film, {cd.pos[0], cd.pos[1], cd.pos[2]}, {cd.look_at[0], cd.look_at[1], cd.look_at[2]},
                                 {cd.up[0], cd.up[1], cd.up[2]}, cd.fovy, cd.film_dist);
        auto sampler = std::make_shared<gy::SobolSampler>();
        auto senum = std::make_shared<gy::SobolEnum>(*film);
        gy::RenderParams params;
        params.rrStartDepth = rp.rr_depth;
        params.mode = GY_RENDER_WAVEFRONT;
        gy::HipPathTracing pt(rp.spp, rp.max_depth, sampler, senum, params, dev);
        pt.Render(scene, camera);

        // the same render straight through the C ABI, in both integrator forms
        gy::DeviceScene ds(*dev, scene);
        gy_camera_desc c2 = camera.Desc();
        gy_render_params p2 = pt.Params();
        std::vector<float> col((size_t)W * H * 3), wgt((size_t)W * H);
        int bad = 0;
        for (int mode : {GY_RENDER_WAVEFRONT, GY_RENDER_MEGAKERNEL}) {
            p2.mode = mode;
            gy::Check(gy_render(ds.Handle(), &c2, &p2, col.data(), wgt.data()), "gy_render");
            if (std::memcmp(col.data(), film->color.data(), col.size() * 4) != 0 ||
                std::memcmp(wgt.data(), film->weight.data(), wgt.size() * 4) != 0) {
                std::printf("film differs from the C-ABI render (mode %d)\n", mode);
                ++bad;
            }
        }
        double sum = 0, wsum = 0;
        for (size_t i = 0; i < wgt.size(); ++i) {
            wsum += film->weight[i];
            for (int k = 0; k < 3; ++k) sum += film->color[3 * i + k];
        }
        const gy::Vec3 c = film->Pixel(W / 2, H / 2);
        // every sample adds weight 1 to its pixel, and to a neighbour too when
        // it lies exactly on a pixel border (visual/Film.h:61-74)
        std::printf("device %s: %dx%d %d spp, weight sum %.0f (at least %d), mean %.6f, centre (%.4f %.4f %.4f), %s\n",
                    dev->Name().c_str(), W, H, rp.spp, wsum, W * H * rp.spp, sum / wsum, c[0], c[1], c[2],
                    bad ? "MISMATCH" : "films identical");
        if (wsum < (double)W * H * rp.spp) ++bad;

        // LightTracing through gy::HipLightTracing: the film's splats equal a
        // direct C-ABI render's up to the float-atomic summation order
        gy::HipLightTracing lt(sampler, 5, 4, params, dev);
        lt.Render(scene, camera);
        gy_render_params p3 = lt.Params();
        std::vector<float> spl(film->splat.size());
        gy::Check(gy_render_light_tracing(ds.Handle(), &c2, &p3, spl.data()), "gy_render_light_tracing");
        double lsum = 0, ldiff = 0;
        for (size_t i = 0; i < spl.size(); ++i) {
            lsum += std::fabs(spl[i]);
            ldiff += std::fabs((double)spl[i] - film->splat[i]);
        }
        const bool lt_ok = lsum > 0 && ldiff <= 1e-5 * lsum;
        std::printf("light tracing: splat sum %.4f, relative difference %.2e, %s\n", lsum, ldiff / (lsum > 0 ? lsum : 1),
                    lt_ok ? "splats agree" : "SPLAT MISMATCH");
        if (!lt_ok) ++bad;
        return bad ? 1 : 0;
    } catch (const gy::Error& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
