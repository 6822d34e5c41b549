// mock_host.cpp -- a C++ stand-in for the Rust host of ppf-contact-solver.
//
// It fills the host's #[repr(C)] DataSet and ParamSet (include/ppf_backend.h)
// from a scene pack (ppf_amd/datapack.py) the way
// crates/ppf-cts-solver/src/builder.rs lays them out, then drives the solver
// backend in the order Backend::run uses (backend.rs:626-1025):
//   set_log_path, initialize, then per step update_constraint (moving pins),
//   refresh_collision_active, advance, fetch, and -- as a checkpointing host
//   does -- fetch_dyn_counts / fetch_dyn and gather_current_positions; on a
//   failed step fetch_intersection_records.
// Output pack: curr / prev after the last step, and per step time, prev_dt,
// the six StepResult flags and the library's cumulative device allocations.
//
//   mock_host <scene.pack> <out.pack> [log_dir]
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ppf_backend.h"
#include "ppf_mi355x.h" // ppf_alloc_counts: the library's allocation counters

namespace {

struct Array {
    char kind = 0;
    std::vector<uint8_t> bytes;
    size_t n = 0;
    template <class T> const T *as() const { return reinterpret_cast<const T *>(bytes.data()); }
};

std::map<std::string, Array> read_pack(const char *path) {
    std::map<std::string, Array> out;
    FILE *fp = fopen(path, "rb");
    if (!fp) {
        fprintf(stderr, "mock_host: cannot open %s\n", path);
        exit(2);
    }
    char magic[8];
    uint32_t count = 0;
    if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, "PPFPACK1", 8) ||
        fread(&count, 4, 1, fp) != 1) {
        fprintf(stderr, "mock_host: %s is not a pack\n", path);
        exit(2);
    }
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t len = 0;
        uint64_t n = 0;
        Array a;
        if (fread(&len, 4, 1, fp) != 1) exit(2);
        std::string name(len, '\0');
        if (fread(&name[0], 1, len, fp) != len || fread(&a.kind, 1, 1, fp) != 1 ||
            fread(&n, 8, 1, fp) != 1)
            exit(2);
        const size_t size = a.kind == 'd' ? 8 : 4;
        a.n = n;
        a.bytes.resize(n * size);
        if (n && fread(a.bytes.data(), size, n, fp) != n) exit(2);
        out[name] = std::move(a);
    }
    fclose(fp);
    return out;
}

void write_array(FILE *fp, const char *name, char kind, const void *data, uint64_t n) {
    const uint32_t len = (uint32_t)strlen(name);
    fwrite(&len, 4, 1, fp);
    fwrite(name, 1, len, fp);
    fwrite(&kind, 1, 1, fp);
    fwrite(&n, 8, 1, fp);
    fwrite(data, kind == 'd' ? 8 : 4, n, fp);
}

struct Pack {
    std::map<std::string, Array> a;
    bool has(const char *k) const { return a.count(k) != 0; }
    const Array &get(const char *k) const {
        auto it = a.find(k);
        if (it == a.end()) {
            fprintf(stderr, "mock_host: pack has no '%s'\n", k);
            exit(2);
        }
        return it->second;
    }
    std::vector<float> f(const char *k) const {
        const Array &x = get(k);
        return std::vector<float>(x.as<float>(), x.as<float>() + x.n);
    }
    std::vector<uint32_t> u(const char *k) const {
        const Array &x = get(k);
        return std::vector<uint32_t>(x.as<uint32_t>(), x.as<uint32_t>() + x.n);
    }
    double d(const char *k) const { return get(k).as<double>()[0]; }
};

template <class V> void set_vec(V &v, void *data, uint32_t size) {
    v.data = static_cast<decltype(v.data)>(data);
    v.size = size;
    v.allocated = size;
}

template <class V> void set_vecvec(V &v, uint32_t *data, uint32_t *offset, uint32_t rows) {
    v.data = data;
    v.offset = offset;
    v.size = rows;
    v.nnz = offset[rows];
    v.nnz_allocated = v.nnz;
    v.offset_allocated = rows + 1;
}

} // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: mock_host <scene.pack> <out.pack> [log_dir]\n");
        return 2;
    }
    Pack pk{read_pack(argv[1])};
    if (argc > 3) {
        set_log_path(argv[3]);
    }
    // ---- geometry and topology (host-owned, as the Rust CVecs are) --------
    std::vector<float> curr = pk.f("x"), prev = pk.f("prev");
    std::vector<uint32_t> edge = pk.u("edge"), face = pk.u("face"), hinge = pk.u("hinge");
    std::vector<uint32_t> v2e_off = pk.u("v2e_off"), v2e = pk.u("v2e"), v2f_off = pk.u("v2f_off"),
                          v2f = pk.u("v2f"), e2f_off = pk.u("e2f_off"), e2f = pk.u("e2f");
    const uint32_t nv = (uint32_t)curr.size() / 3, ne = (uint32_t)edge.size() / 2,
                   nf = (uint32_t)face.size() / 3, nh = (uint32_t)hinge.size() / 4;
    cts_DataSet ds{};
    set_vec(ds.vertex.curr, curr.data(), nv);
    set_vec(ds.vertex.prev, prev.data(), nv);
    set_vec(ds.mesh.mesh.face, face.data(), nf);
    set_vec(ds.mesh.mesh.edge, edge.data(), ne);
    set_vec(ds.mesh.mesh.hinge, hinge.data(), nh);
    set_vecvec(ds.mesh.neighbor.vertex.face, v2f.data(), v2f_off.data(), nv);
    set_vecvec(ds.mesh.neighbor.vertex.edge, v2e.data(), v2e_off.data(), nv);
    set_vecvec(ds.mesh.neighbor.edge.face, e2f.data(), e2f_off.data(), ne);

    // ---- props and one parameter record per element ------------------------
    const std::vector<float> v_mass = pk.f("v_mass"), v_area = pk.f("v_area"),
                             v_ghat = pk.f("v_ghat"), v_off = pk.f("v_off"),
                             v_fric = pk.f("v_fric");
    const std::vector<uint32_t> v_col = pk.u("v_collider");
    std::vector<cts_VertexProp> vprop(nv);
    std::vector<cts_VertexParam> vparam(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        vprop[i] = cts_VertexProp{};
        vprop[i].area = v_area[i];
        vprop[i].mass = v_mass[i];
        vprop[i].param_index = i;
        vprop[i].collider = v_col[i] != 0;
        vparam[i] = cts_VertexParam{v_ghat[i], v_off[i], v_fric[i]};
    }
    const std::vector<float> e_mass = pk.f("e_mass"), e_len = pk.f("e_length"),
                             e_ghat = pk.f("e_ghat"), e_off = pk.f("e_off"),
                             e_fric = pk.f("e_fric");
    std::vector<cts_EdgeProp> eprop(ne);
    std::vector<cts_EdgeParam> eparam(ne);
    for (uint32_t e = 0; e < ne; ++e) {
        eprop[e] = cts_EdgeProp{};
        eprop[e].length = eprop[e].initial_length = e_len[e];
        eprop[e].mass = e_mass[e];
        eprop[e].param_index = e;
        eparam[e] = cts_EdgeParam{};
        eparam[e].ghat = e_ghat[e];
        eparam[e].offset = e_off[e];
        eparam[e].friction = e_fric[e];
    }
    const std::vector<float> f_mass = pk.f("f_mass"), f_mu = pk.f("f_mu"),
                             f_lam = pk.f("f_lambda"), f_ghat = pk.f("f_ghat"),
                             f_off = pk.f("f_off"), f_fric = pk.f("f_fric");
    const std::vector<float> f_sl = pk.has("f_strainlimit") ? pk.f("f_strainlimit")
                                                            : std::vector<float>(nf, 0.0f);
    const std::vector<float> f_pl = pk.has("f_plasticity") ? pk.f("f_plasticity")
                                                           : std::vector<float>(nf, 0.0f);
    const std::vector<float> f_plt = pk.has("f_plastic_threshold")
                                         ? pk.f("f_plastic_threshold")
                                         : std::vector<float>(nf, 0.0f);
    std::vector<cts_FaceProp> fprop(nf);
    std::vector<cts_FaceParam> fparam(nf);
    for (uint32_t f = 0; f < nf; ++f) {
        fprop[f] = cts_FaceProp{};
        fprop[f].mass = f_mass[f];
        fprop[f].param_index = f;
        cts_FaceParam &p = fparam[f];
        p = cts_FaceParam{};
        p.model = CTS_MODEL_BARAFF_WITKIN;
        p.mu = f_mu[f], p.lambda = f_lam[f], p.friction = f_fric[f];
        p.ghat = f_ghat[f], p.offset = f_off[f];
        p.shrink_x = p.shrink_y = 1.0f;
        p.strainlimit = f_sl[f];
        p.plasticity = f_pl[f], p.plasticity_threshold = f_plt[f];
    }
    const std::vector<float> h_rest = pk.f("h_rest"), h_len = pk.f("h_length"),
                             h_area = pk.f("h_area"), h_bend = pk.f("h_bend");
    const std::vector<float> h_pl = pk.has("h_plasticity") ? pk.f("h_plasticity")
                                                           : std::vector<float>(nh, 0.0f);
    const std::vector<float> h_plt = pk.has("h_plastic_threshold")
                                         ? pk.f("h_plastic_threshold")
                                         : std::vector<float>(nh, 0.0f);
    std::vector<cts_HingeProp> hprop(nh);
    std::vector<cts_HingeParam> hparam(nh);
    for (uint32_t h = 0; h < nh; ++h) {
        hprop[h] = cts_HingeProp{};
        hprop[h].length = h_len[h], hprop[h].area = h_area[h], hprop[h].rest_angle = h_rest[h];
        hprop[h].uv_edge_sin2 = -1.0f;
        hprop[h].param_index = h;
        hparam[h] = cts_HingeParam{};
        hparam[h].bend = h_bend[h];
        hparam[h].plasticity = h_pl[h], hparam[h].plasticity_threshold = h_plt[h];
    }
    set_vec(ds.prop.vertex, vprop.data(), nv);
    set_vec(ds.prop.edge, eprop.data(), ne);
    set_vec(ds.prop.face, fprop.data(), nf);
    set_vec(ds.prop.hinge, hprop.data(), nh);
    set_vec(ds.param_arrays.vertex, vparam.data(), nv);
    set_vec(ds.param_arrays.edge, eparam.data(), ne);
    set_vec(ds.param_arrays.face, fparam.data(), nf);
    set_vec(ds.param_arrays.hinge, hparam.data(), nh);
    std::vector<float> inv_rest = pk.f("inv_rest");
    set_vec(ds.inv_rest2x2, inv_rest.data(), nf);
    ds.shell_face_count = nf;
    ds.surface_vert_count = nv;

    // ---- pins --------------------------------------------------------------
    const std::vector<uint32_t> pin_index = pk.u("pin_index");
    const uint32_t npin = (uint32_t)pin_index.size();
    std::vector<cts_FixPair> fix(npin);
    auto set_pins = [&](const float *pos, const float *delta) {
        for (uint32_t k = 0; k < npin; ++k) {
            fix[k] = cts_FixPair{};
            for (int d = 0; d < 3; ++d) {
                fix[k].position.v[d] = pos[3 * k + d];
                fix[k].step_delta.v[d] = delta[3 * k + d];
            }
            fix[k].index = pin_index[k];
            fix[k].kinematic = true;
        }
    };
    if (npin) {
        const std::vector<float> pos = pk.f("pin_pos"), del = pk.f("pin_delta");
        set_pins(pos.data(), del.data());
    }
    set_vec(ds.constraint.fix, fix.data(), npin);

    // ---- static collision mesh ----------------------------------------------
    std::vector<float> s_x;
    std::vector<uint32_t> s_edge, s_face;
    std::vector<cts_VertexProp> s_vprop;
    std::vector<cts_EdgeProp> s_eprop;
    std::vector<cts_FaceProp> s_fprop;
    std::vector<cts_VertexParam> s_vparam;
    std::vector<cts_EdgeParam> s_eparam;
    std::vector<cts_FaceParam> s_fparam;
    if (pk.has("s_x")) {
        s_x = pk.f("s_x"), s_edge = pk.u("s_edge"), s_face = pk.u("s_face");
        const uint32_t snv = (uint32_t)s_x.size() / 3, sne = (uint32_t)s_edge.size() / 2,
                       snf = (uint32_t)s_face.size() / 3;
        const std::vector<float> vg = pk.f("s_v_ghat"), vo = pk.f("s_v_off"),
                                 vf = pk.f("s_v_fric"), eg = pk.f("s_e_ghat"),
                                 eo = pk.f("s_e_off"), ef = pk.f("s_e_fric"),
                                 fg = pk.f("s_f_ghat"), fo = pk.f("s_f_off"),
                                 ff = pk.f("s_f_fric");
        s_vprop.assign(snv, cts_VertexProp{});
        s_vparam.resize(snv);
        for (uint32_t i = 0; i < snv; ++i) {
            s_vprop[i].param_index = i;
            s_vparam[i] = cts_VertexParam{vg[i], vo[i], vf[i]};
        }
        s_eprop.assign(sne, cts_EdgeProp{});
        s_eparam.assign(sne, cts_EdgeParam{});
        for (uint32_t e = 0; e < sne; ++e) {
            s_eprop[e].param_index = e;
            s_eparam[e].ghat = eg[e], s_eparam[e].offset = eo[e], s_eparam[e].friction = ef[e];
        }
        s_fprop.assign(snf, cts_FaceProp{});
        s_fparam.assign(snf, cts_FaceParam{});
        for (uint32_t f = 0; f < snf; ++f) {
            s_fprop[f].param_index = f;
            s_fparam[f].model = CTS_MODEL_BARAFF_WITKIN;
            s_fparam[f].ghat = fg[f], s_fparam[f].offset = fo[f], s_fparam[f].friction = ff[f];
            s_fparam[f].shrink_x = s_fparam[f].shrink_y = 1.0f;
        }
        cts_CollisionMesh &cm = ds.constraint.mesh;
        set_vec(cm.vertex, s_x.data(), snv);
        set_vec(cm.edge, s_edge.data(), sne);
        set_vec(cm.face, s_face.data(), snf);
        set_vec(cm.prop.vertex, s_vprop.data(), snv);
        set_vec(cm.prop.edge, s_eprop.data(), sne);
        set_vec(cm.prop.face, s_fprop.data(), snf);
        set_vec(cm.param_arrays.vertex, s_vparam.data(), snv);
        set_vec(cm.param_arrays.edge, s_eparam.data(), sne);
        set_vec(cm.param_arrays.face, s_fparam.data(), snf);
    }

    // Analytic colliders: rows of 10 floats -- sphere: center xyz, ghat,
    // friction, radius, thickness, bowl, reverse, kinematic; floor: ground
    // xyz, ghat, friction, thickness, up xyz, kinematic.
    std::vector<cts_Sphere> spheres;
    std::vector<cts_Floor> floors;
    if (pk.has("spheres")) {
        const std::vector<float> a = pk.f("spheres");
        for (size_t k = 0; k + 10 <= a.size(); k += 10) {
            cts_Sphere sp{};
            for (int d = 0; d < 3; ++d) sp.center.v[d] = a[k + d];
            sp.ghat = a[k + 3], sp.friction = a[k + 4], sp.radius = a[k + 5];
            sp.thickness = a[k + 6];
            sp.bowl = a[k + 7] != 0.0f, sp.reverse = a[k + 8] != 0.0f;
            sp.kinematic = a[k + 9] != 0.0f;
            spheres.push_back(sp);
        }
        set_vec(ds.constraint.sphere, spheres.data(), (uint32_t)spheres.size());
    }
    if (pk.has("floors")) {
        const std::vector<float> a = pk.f("floors");
        for (size_t k = 0; k + 10 <= a.size(); k += 10) {
            cts_Floor fl{};
            for (int d = 0; d < 3; ++d) fl.ground.v[d] = a[k + d];
            fl.ghat = a[k + 3], fl.friction = a[k + 4], fl.thickness = a[k + 5];
            for (int d = 0; d < 3; ++d) fl.up.v[d] = a[k + 6 + d];
            fl.kinematic = a[k + 9] != 0.0f;
            floors.push_back(fl);
        }
        set_vec(ds.constraint.floor, floors.data(), (uint32_t)floors.size());
    }
    // A stitch, which this backend refuses (initialize must fail with fatal
    // code 4 and name it).
    cts_Stitch stitch{};
    if (pk.has("stitch")) {
        set_vec(ds.constraint.stitch, &stitch, 1);
    }

    // ---- ParamSet -------------------------------------------------------------
    cts_ParamSet param{};
    param.time = 0.0;
    param.dt = (float)pk.d("dt");
    param.prev_dt = (float)pk.d("prev_dt");
    param.playback = 1.0f;
    param.cg_tol = (float)pk.d("cg_tol");
    param.constraint_tol = pk.has("constraint_tol") ? (float)pk.d("constraint_tol") : 0.01f;
    param.cg_max_iter = (uint32_t)pk.d("cg_max_iter");
    param.eiganalysis_eps = pk.has("eiganalysis_eps") ? (float)pk.d("eiganalysis_eps") : 1e-2f;
    param.target_toi = (float)pk.d("target_toi");
    param.line_search_max_t = (float)pk.d("line_search_max_t");
    param.ccd_eps = (float)pk.d("ccd_eps");
    param.max_dx = (float)pk.d("max_dx");
    param.friction_eps = (float)pk.d("friction_eps");
    param.friction_mode = (uint32_t)pk.d("friction_mode");
    param.min_newton_steps = (uint32_t)pk.d("min_newton_steps");
    param.max_newton_steps = (uint32_t)pk.d("max_newton_steps");
    param.air_density = (float)pk.d("air_density");
    param.air_friction = (float)pk.d("air_friction");
    param.disable_contact = pk.d("disable_contact") != 0.0;
    param.barrier = CTS_BARRIER_CUBIC;
    param.precond = CTS_PRECOND_BLOCK_JACOBI;
    const std::vector<float> grav = pk.f("gravity"), wind = pk.f("wind");
    for (int d = 0; d < 3; ++d) {
        param.gravity.v[d] = grav[d];
        param.wind.v[d] = wind[d];
    }

    // ---- Backend::run's order -----------------------------------------------
    if (!initialize(&ds, &param)) {
        fprintf(stderr, "mock_host: initialize failed (code %u): %s\n", ppf_fatal_code(),
                ppf_fatal_detail());
        return 3;
    }
    if (pk.has("cw_dmap")) {
        const std::vector<uint32_t> dmap = pk.u("cw_dmap"), counts = pk.u("cw_counts");
        const std::vector<float> win = pk.f("cw_windows");
        init_collision_windows(dmap.data(), (uint32_t)dmap.size(), win.data(), counts.data(),
                               (uint32_t)counts.size());
    }
    // A resumed run hands the checkpoint's contact pattern back (update_dyn);
    // what fetch_dyn reports before the first step is written out.
    std::vector<uint32_t> dyn0_off, dyn0_idx;
    if (pk.has("carry_dyn_offset")) {
        const std::vector<uint32_t> off = pk.u("carry_dyn_offset"), idx = pk.u("carry_dyn_index");
        update_dyn(idx.data(), off.data());
        uint32_t nval = 0, noff = 0;
        fetch_dyn_counts(&nval, &noff);
        dyn0_off.resize(noff);
        dyn0_idx.resize(nval);
        fetch_dyn(dyn0_idx.data(), nullptr, dyn0_off.data());
    }
    const uint32_t steps = pk.u("steps")[0];
    std::vector<float> sched, dsched;
    if (pk.has("pin_schedule")) {
        sched = pk.f("pin_schedule");
        dsched = pk.f("pin_delta_schedule");
    }
    std::vector<double> times;
    std::vector<float> dts;
    std::vector<uint32_t> flags, allocs, frees, dyn_off, dyn_idx;
    std::vector<uint32_t> tracked;
    for (uint32_t v = 0; v < nv && tracked.size() < 16; v += 1 + nv / 16) {
        tracked.push_back(v);
    }
    std::vector<float> gathered(3 * tracked.size());
    int rc = 0;
    for (uint32_t k = 0; k < steps; ++k) {
        if (!sched.empty()) {
            set_pins(sched.data() + (size_t)k * npin * 3, dsched.data() + (size_t)k * npin * 3);
            update_constraint(&ds.constraint);
        }
        refresh_collision_active((float)param.time);
        cts_StepResult r{};
        advance(&r);
        flags.insert(flags.end(), {r.ccd_success, r.pcg_success, r.intersection_free,
                                   r.newton_progress, r.pin_feasible, r.contact_separated});
        times.push_back(r.time);
        dts.push_back(param.prev_dt);
        const bool ok = r.ccd_success && r.pcg_success && r.intersection_free &&
                        r.newton_progress && r.pin_feasible && r.contact_separated;
        if (!ok) {
            cts_IntersectionRecord rec[8];
            const uint32_t n = fetch_intersection_records(rec, 8);
            fprintf(stderr, "mock_host: step %u failed (%u intersection records)\n", k, n);
            rc = 4;
            break;
        }
        fetch();
        // a checkpointing host: the contact pattern (state_<N>.bin.gz,
        // backend.rs:520-575) and a few tracked vertices every step
        uint32_t nval = 0, noff = 0;
        fetch_dyn_counts(&nval, &noff);
        dyn_off.resize(noff);
        dyn_idx.resize(nval);
        fetch_dyn(dyn_idx.data(), nullptr, dyn_off.data());
        gather_current_positions(tracked.data(), (uint32_t)tracked.size(), gathered.data());
        uint64_t na = 0, nf = 0;
        ppf_alloc_counts(&na, &nf);
        allocs.push_back((uint32_t)na);
        frees.push_back((uint32_t)nf);
        printf("step %u: time %.6f dt %.6g\n", k, param.time, param.prev_dt);
        fflush(stdout);
    }
    uint32_t n_value = 0, n_offset = 0;
    fetch_dyn_counts(&n_value, &n_offset);
    // the rest shape as the run left it (plasticity), as a checkpointing host
    // pulls it
    fetch_inv_rest();
    fetch_rest_angles();
    std::vector<float> rest_angle(nh);
    for (uint32_t h = 0; h < nh; ++h) {
        rest_angle[h] = hprop[h].rest_angle;
    }
    FILE *fp = fopen(argv[2], "wb");
    if (!fp) {
        return 2;
    }
    const uint32_t count = 12;
    fwrite("PPFPACK1", 1, 8, fp);
    fwrite(&count, 4, 1, fp);
    write_array(fp, "curr", 'f', curr.data(), curr.size());
    write_array(fp, "prev", 'f', prev.data(), prev.size());
    write_array(fp, "time", 'd', times.data(), times.size());
    write_array(fp, "dt", 'f', dts.data(), dts.size());
    write_array(fp, "flags", 'u', flags.data(), flags.size());
    const uint32_t dyn[2] = {n_value, n_offset};
    write_array(fp, "dyn_counts", 'u', dyn, 2);
    // device allocations / frees of the library after each step (cumulative)
    write_array(fp, "allocs", 'u', allocs.data(), allocs.size());
    write_array(fp, "frees", 'u', frees.data(), frees.size());
    write_array(fp, "dyn0_offset", 'u', dyn0_off.data(), dyn0_off.size());
    write_array(fp, "dyn0_index", 'u', dyn0_idx.data(), dyn0_idx.size());
    write_array(fp, "inv_rest", 'f', inv_rest.data(), inv_rest.size());
    write_array(fp, "rest_angle", 'f', rest_angle.data(), rest_angle.size());
    fclose(fp);
    return rc;
}
