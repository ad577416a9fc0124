// Renders the Cornell box through gy::HipPathTracing (the Integrator
// interface) and checks the film against direct C-ABI renders: the adapter
// must add nothing, and the megakernel and wavefront films must be equal bit
// for bit.  Then the same for gy::HipLightTracing, gy::HipVolPathTracing (the
// room with a slab of medium) and gy::HipSPPM.  Exit status 0 on success.
//
//   render_cornell [--expect-no-device]
// With --expect-no-device the program checks that opening a device fails with
// a gy::Error (hosts without a GPU) and exits 0.
#include <gyunity/HipPathTracing.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>

// The reference-side flattening of GpuPathTracing::Flatten,
// run over the generated scene's arrays, which stand in for
// Scene::GetPrimitives() and Scene::GetLights(): primitives in insertion
// order with one gy::Material per distinct material (numbered in first-use
// order), then the area lights in light order.
static gy::Scene FlattenPrimitives(const gy_scene_desc& d) {
    gy::Scene out;
    std::map<uint32_t, int> mat_ids;
    for (int i = 0; i < d.n_tri; ++i) {
        const uint32_t m = d.mat[i];
        auto it = mat_ids.find(m);
        int mid;
        if (it != mat_ids.end()) {
            mid = it->second;
        } else {
            gy::Material gm;
            std::copy(d.mats + (size_t)m * GY_MAT_WORDS, d.mats + (size_t)(m + 1) * GY_MAT_WORDS, gm.words.begin());
            mid = mat_ids[m] = out.AddMaterial(gm);
        }
        const float* p = d.pos + 9 * (size_t)i;
        const float* n = d.nrm + 9 * (size_t)i;
        const float* f = d.face_n + 3 * (size_t)i;
        std::array<float, 6> uv = {0, 0, 0, 0, 0, 0};
        if (d.uv) std::copy(d.uv + 6 * (size_t)i, d.uv + 6 * (size_t)i + 6, uv.begin());
        out.AddTriangle({gy::Vec3{p[0], p[1], p[2]}, gy::Vec3{p[3], p[4], p[5]}, gy::Vec3{p[6], p[7], p[8]}},
                        {gy::Vec3{n[0], n[1], n[2]}, gy::Vec3{n[3], n[4], n[5]}, gy::Vec3{n[6], n[7], n[8]}},
                        gy::Vec3{f[0], f[1], f[2]}, mid, uv);
    }
    for (int l = 0; l < d.n_light; ++l)
        out.AddAreaLight(d.light_tri[l], gy::Vec3{d.light_le[3 * l], d.light_le[3 * l + 1], d.light_le[3 * l + 2]});
    return out;
}

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
        gy_render_params rp = {};
        gy::Check(gy_generated_desc(g, &desc, &cd, &rp), "gy_generated_desc");
        const gy::Scene scene = gy::Scene::FromDesc(desc);
        const gy::Scene flat = FlattenPrimitives(desc);
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
        // the scene flattened primitive by primitive renders the same film
        {
            auto film2 = std::make_shared<gy::Film>(W, H);
            gy::PinHoleCamera camera2(film2, {cd.pos[0], cd.pos[1], cd.pos[2]}, {cd.look_at[0], cd.look_at[1], cd.look_at[2]},
                                      {cd.up[0], cd.up[1], cd.up[2]}, cd.fovy, cd.film_dist);
            gy::HipPathTracing pt2(rp.spp, rp.max_depth, sampler, senum, params, dev);
            pt2.Render(flat, camera2);
            const bool same = film2->color == film->color && film2->weight == film->weight;
            std::printf("primitive-by-primitive flattening: %s\n", same ? "same film" : "FILM DIFFERS");
            if (!same) ++bad;
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

        // A copy of the scene with an analytic sphere is another scene (new
        // stamp): the adapter uploads it, its film equals the C-ABI render of
        // it, and rendering the first scene again gives the first film back.
        {
            const std::vector<float> first_color = film->color, first_weight = film->weight;
            gy::Scene with_sphere = scene;
            with_sphere.AddSphere({0.3f, 0.18f, 0.55f}, 0.15f, 0);
            pt.Render(with_sphere, camera);
            gy::DeviceScene dss(*dev, with_sphere);
            p2.mode = GY_RENDER_MEGAKERNEL;
            gy::Check(gy_render(dss.Handle(), &c2, &p2, col.data(), wgt.data()), "gy_render");
            const bool sph_same = std::memcmp(col.data(), film->color.data(), col.size() * 4) == 0 &&
                                  std::memcmp(wgt.data(), film->weight.data(), wgt.size() * 4) == 0;
            const bool sph_differs = film->color != first_color;
            pt.Render(scene, camera);
            const bool back = film->color == first_color && film->weight == first_weight;
            std::printf("scene with a sphere: %s, %s; first scene again: %s\n", sph_same ? "same film as the C ABI" : "FILM DIFFERS",
                        sph_differs ? "differs from the first" : "SAME AS THE FIRST", back ? "first film" : "FILM DIFFERS");
            if (!sph_same || !sph_differs || !back) ++bad;
        }

        // VolPathTracing through gy::HipVolPathTracing on the room with a
        // material-less slab of medium (12 outward-facing triangles, medium
        // inside, none outside): the film equals direct C-ABI renders in both
        // forms.  SPPM through gy::HipSPPM: the image equals a direct call.
        {
            gy::Scene fog = scene;
            const int medium = fog.AddMedium({0.3f, 0.6f, 0.9f}, {2.0f, 1.5f, 1.0f}, 0.5f);
            const float lo[3] = {0.15f, 0.62f, 0.1f}, hi[3] = {0.85f, 0.92f, 0.9f};
            for (int axis = 0; axis < 3; ++axis)
                for (int side = 0; side < 2; ++side) {
                    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
                    gy::Vec3 n = {0, 0, 0};
                    n[axis] = side ? 1.f : -1.f;
                    std::array<gy::Vec3, 4> q;
                    for (int c = 0; c < 4; ++c) {
                        q[c][axis] = side ? hi[axis] : lo[axis];
                        q[c][u] = (c == 1 || c == 2) ? hi[u] : lo[u];
                        q[c][v] = (c >= 2) ? hi[v] : lo[v];
                    }
                    for (const auto& t : {std::array<int, 3>{0, 1, 2}, std::array<int, 3>{0, 2, 3}}) {
                        const int id = fog.AddTriangle({q[t[0]], q[t[1]], q[t[2]]}, {n, n, n}, n, 0);
                        fog.SetTriangleMedium(id, medium, -1, true);
                    }
                }
            gy::RenderParams vparams;
            vparams.mode = GY_RENDER_WAVEFRONT;
            gy::HipVolPathTracing vpt(4, 6, sampler, senum, true, vparams, dev);
            vpt.Render(fog, camera);
            gy::DeviceScene dsf(*dev, fog);
            gy_render_params pv = vpt.Params();
            bool vpt_same = true;
            for (int mode : {GY_RENDER_WAVEFRONT, GY_RENDER_MEGAKERNEL}) {
                pv.mode = mode;
                gy::Check(gy_render_vpt(dsf.Handle(), &c2, &pv, 1, col.data(), wgt.data()), "gy_render_vpt");
                vpt_same = vpt_same && std::memcmp(col.data(), film->color.data(), col.size() * 4) == 0 &&
                           std::memcmp(wgt.data(), film->weight.data(), wgt.size() * 4) == 0;
            }
            double vsum = 0;
            for (float x : film->color) vsum += x;
            std::printf("volumetric path tracing: colour sum %.2f, %s\n", vsum, vpt_same ? "same film as the C ABI" : "FILM DIFFERS");
            if (!vpt_same || !(vsum > 0)) ++bad;

            gy::HipSPPM sppm(2, 20000, 5, 0.05f, 0.7f, false, sampler, senum, dev);
            sppm.Render(scene, camera);
            gy_sppm_params ps = sppm.Params();
            gy::Check(gy_render_sppm(ds.Handle(), &c2, &ps, col.data(), nullptr, nullptr), "gy_render_sppm");
            const bool sppm_same = std::memcmp(col.data(), film->color.data(), col.size() * 4) == 0;
            const gy::Vec3 px = film->Pixel(W / 2, H / 2);
            std::printf("sppm: centre (%.4f %.4f %.4f), %s\n", px[0], px[1], px[2], sppm_same ? "same image as the C ABI" : "IMAGE DIFFERS");
            if (!sppm_same) ++bad;
        }
        return bad ? 1 : 0;
    } catch (const gy::Error& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
