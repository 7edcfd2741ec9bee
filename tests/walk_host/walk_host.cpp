// Host build of the device prism walk (haseongpu_amd/csrc/prism_walk.h) for
// CPU tests: the same source, compiled for the host, walked over meshes the
// test packs into the device record layouts (48-byte global records and the
// 32-byte records staged in LDS).  Test infrastructure only.
#include "../../haseongpu_amd/csrc/prism_walk.h"

using namespace hase;

extern "C" int walk_rays_host(uint32_t T, uint32_t layers, float h, const uint32_t* rec, const uint32_t* lds_rec,
                              const float* beta, const float* refl, const float* nidx, uint32_t n,
                              const float* start, const uint32_t* prism, const float* end, const uint32_t* nrefl,
                              const uint32_t* first, float* out, uint32_t* steps, int32_t* status)
{
    MeshDev m{};
    m.rec = reinterpret_cast<const uint4*>(rec);
    m.beta = beta;
    m.refl = refl;
    m.T = T;
    m.layers = layers;
    m.h = h;
    if (nidx) {
        m.has_nidx = 1;
        m.nin_b = nidx[0]; m.nout_b = nidx[1]; m.nin_t = nidx[2]; m.nout_t = nidx[3];
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t t = prism[i] % T, l = prism[i] / T;
        WalkOut w;
        if (lds_rec)
            w = walk_ray(m, start[3 * i], start[3 * i + 1], start[3 * i + 2], t, l, end[3 * i], end[3 * i + 1],
                         end[3 * i + 2], nrefl[i], first[i], LdsTriangles{reinterpret_cast<const uint4*>(lds_rec)});
        else
            w = walk_ray(m, start[3 * i], start[3 * i + 1], start[3 * i + 2], t, l, end[3 * i], end[3 * i + 1],
                         end[3 * i + 2], nrefl[i], first[i]);
        out[5 * i + 0] = w.beta_len;
        out[5 * i + 1] = w.gain_len;
        out[5 * i + 2] = w.clad_len;
        out[5 * i + 3] = w.refl;
        out[5 * i + 4] = w.length;
        steps[i] = w.steps;
        status[i] = w.status;
    }
    return 0;
}
