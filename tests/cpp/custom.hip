/*
 * User-defined device integrands through the drop-in header (viltrum_amd/device.hpp),
 * written the way a user of the reference writes an integrand: a functor or a lambda
 * in each of the adapter's call forms (src/integrate.h:11-70).  Compiled by hipcc for
 * gfx950 and linked to the in-tree libviltrum_amd.so.
 *
 * Usage: custom <out_dir>.  Writes the bins of each case as raw arrays into out_dir and
 * prints one line per case; tests/test_gpu_custom.py checks them.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "viltrum_amd/device.hpp"

namespace {

/* The same function as the built-in polynomial family (vmx::Poly), in the same order
   of operations: terms in index order, powers by repeated multiplication, then the
   half-space step. */
template <int D>
struct PolyFn {
    int nterms;
    double coef[4];
    int expo[4][D];
    double step_n[D], step_d, step_h;

    template <typename T>
    __host__ __device__ T operator()(const T* x) const {
        T s = T(0);
        for (int m = 0; m < nterms; ++m) {
            T t = T(coef[m]);
            for (int d = 0; d < D; ++d)
                for (int e = 0; e < expo[m][d]; ++e) t = t * x[d];
            s = s + t;
        }
        T lin = T(0);
        for (int d = 0; d < D; ++d) lin = lin + T(step_n[d]) * x[d];
        if (lin > T(step_d)) s = s + T(step_h);
        return s;
    }
};

template <int D>
vmx_poly_params poly_params(const PolyFn<D>& f) {
    vmx_poly_params p;
    std::memset(&p, 0, sizeof(p));
    p.nterms = f.nterms;
    p.has_step = 1;
    for (int m = 0; m < f.nterms; ++m) {
        p.coef[m] = f.coef[m];
        for (int d = 0; d < D; ++d) p.expo[m][d] = uint8_t(f.expo[m][d]);
    }
    for (int d = 0; d < D; ++d) p.step_n[d] = f.step_n[d];
    p.step_d = f.step_d;
    p.step_h = f.step_h;
    return p;
}

template <typename T>
struct Grid {
    std::vector<T> v;
    std::size_t r0, r1, r2;
    Grid(std::size_t a, std::size_t b, std::size_t c = 1) : v(a * b * c, T(0)), r0(a), r1(b), r2(c) {}
    T& operator()(const std::array<std::size_t, 2>& i) { return v[i[0] * r1 + i[1]]; }
    T& operator()(const std::array<std::size_t, 3>& i) { return v[(i[0] * r1 + i[1]) * r2 + i[2]]; }
};

template <typename T>
void save(const std::string& dir, const std::string& name, const std::vector<T>& v) {
    const std::string path = dir + "/" + name + ".bin";
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp || std::fwrite(v.data(), sizeof(T), v.size(), fp) != v.size()) throw std::runtime_error("write " + path);
    std::fclose(fp);
}

template <typename T>
bool same_bits(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(T) * a.size()) == 0;
}

/* The built-in Dyson series (vmx::Dyson) written against the sequence a user integrand
   of an infinite range reads, in the same order of operations. */
struct DysonFn {
    float decay;
    template <class Seq>
    __host__ __device__ float operator()(const Seq& seq) const {
        auto it = seq.begin();
        float sum = 1.f, range = 1.f;
        for (int step = 0; step < 4096 && *it < decay; ++step) {
            ++it;
            const float x = range * (*it);
            const float prob = float(1.0 / double(range));
            range = x;
            sum += x / prob;
            ++it;
        }
        return sum;
    }
};

/* The reference's documented integrand of an infinite range (main/doc/ranges-infinite.cc:
   5-27): a series cut by Russian roulette on every other entry.  Over the primary
   infinite range with one region dim, E[f | x0] = [x0 < d] / (1 - d). */
struct SeriesFn {
    float decaying_factor;
    template <class Seq>
    __host__ __device__ float operator()(const Seq& seq) const {
        auto x = seq.begin();
        float sum = 0.0f, term = 1.0f;
        while ((*x) < decaying_factor) {
            ++x;
            term *= 2.0f * (*x);
            ++x;
            sum += term;
        }
        return sum;
    }
};

void hip_ok(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: custom <out_dir>\n");
        return 2;
    }
    const std::string out = argv[1];
    namespace vi = viltrum_amd;

    // 1. a functor equal to a built-in polynomial: the same bins bit for bit, in FP64
    //    (reference region order) and FP32, at 2 and 4 dims (4: the region-major kernel)
    const PolyFn<2> p2{3, {1.0, 0.5, 0.25, 0.0}, {{1, 2}, {2, 0}, {0, 0}, {0, 0}}, {1.0, 1.0}, 1.0, 0.5};
    const auto f2 = vi::device_integrand<2, float, double>(p2);
    const auto b2 = vi::poly(poly_params(p2), 2);
    {
        const vi::IntegratorGpu it(300, 4, 11, VMX_HEUR_SIZE_RELATIVE, VMX_FORM_VR_OPTIMIZE, 1.0, VMX_ORDER_REFERENCE);
        Grid<double> a(24, 20), b(24, 20);
        vi::integrate(it, a, std::array<std::size_t, 2>{24, 20}, f2, vi::range_primary<2>());
        vi::integrate(it, b, std::array<std::size_t, 2>{24, 20}, b2, vi::range_primary<2>());
        std::printf("poly2_f64 %d\n", int(same_bits(a.v, b.v)));
    }
    {
        const auto it = vi::integrator_crespo2021(400, 16, 3);
        Grid<float> a(32, 32), b(32, 32);
        vi::integrate(it, a, std::array<std::size_t, 2>{32, 32}, f2, vi::range_primary<2>());
        vi::integrate(it, b, std::array<std::size_t, 2>{32, 32}, b2, vi::range_primary<2>());
        std::printf("poly2_f32 %d\n", int(same_bits(a.v, b.v)));
    }
    {
        const PolyFn<4> p4{3, {1.0, 0.5, 0.25, 0.0}, {{1, 2, 0, 3}, {2, 0, 1, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
                           {1.0, 1.0, 1.0, 1.0}, 2.0, 0.5};
        const auto it = vi::integrator_crespo2021(200, 5, 7);
        Grid<float> a(9, 7), b(9, 7);
        vi::integrate(it, a, std::array<std::size_t, 2>{9, 7}, vi::device_integrand<4>(p4), vi::range_primary<4>());
        vi::integrate(it, b, std::array<std::size_t, 2>{9, 7}, vi::poly(poly_params(p4), 4), vi::range_primary<4>());
        std::printf("poly4_f32 %d\n", int(same_bits(a.v, b.v)));
    }

    // 2. a lambda f(x, y), tensor-quadratic: the control variate is exact (FP64 bins)
    const auto quad = [] __host__ __device__(double x, double y) {
        return (1.0 + x - 0.5 * x * x) * (2.0 - y + 3.0 * y * y);
    };
    {
        std::vector<std::vector<double>> bins(16, std::vector<double>(16));
        vi::integrate(vi::integrator_crespo2021(63, 8, 5), bins, vi::device_integrand<2, double>(quad),
                      vi::range_primary<2>());
        std::vector<double> flat;
        for (const auto& r : bins) flat.insert(flat.end(), r.begin(), r.end());
        save(out, "quad2_f64", flat);
        std::printf("quad2_f64 %zu\n", flat.size());
    }

    // 3. a transcendental lambda in FP32 (closed-form bins on the Python side)
    const auto trans = [] __host__ __device__(float x, float y) { return sinf(3.f * x) * expf(y); };
    const auto ft = vi::device_integrand<2>(trans);
    {
        std::vector<std::vector<float>> bins(32, std::vector<float>(32));
        vi::integrate(vi::integrator_crespo2021(1023, 16, 1), bins, ft, vi::range_primary<2>());
        std::vector<float> flat;
        for (const auto& r : bins) flat.insert(flat.end(), r.begin(), r.end());
        save(out, "trans2_f32", flat);
        std::printf("trans2_f32 %zu\n", flat.size());
    }

    // 4. f(std::array<T, 3>) over three dims with two bin dims
    struct Arr3 {
        float w;
        __host__ __device__ float operator()(const std::array<float, 3>& a) const { return a[0] * a[1] + w * a[2]; }
    };
    {
        Grid<float> g(12, 10);
        vi::integrate(vi::integrator_crespo2021(200, 8, 2), g, std::array<std::size_t, 2>{12, 10},
                      vi::device_integrand<3>(Arr3{0.75f}), vi::range_primary<3>());
        save(out, "arr3_f32", g.v);
        std::printf("arr3_f32 %zu\n", g.v.size());
    }

    // 5. one batch of three integrals of one functor type, different captured
    //    parameters: each gives what its single call gives, bit for bit
    {
        struct Scaled {
            float a, b;
            __host__ __device__ float operator()(const float* x) const {
                return a * expf(-b * ((x[0] - 0.4f) * (x[0] - 0.4f) + (x[1] - 0.6f) * (x[1] - 0.6f)));
            }
        };
        const Scaled ps[3] = {{1.f, 20.f}, {2.f, 5.f}, {0.5f, 60.f}};
        std::vector<vi::Integrand> fs;
        for (const auto& s : ps) fs.push_back(vi::device_integrand<2>(s));
        vi::Context& ctx = vi::Context::thread_default();
        std::vector<vmx_problem> probs(3);
        std::vector<vmx_integrand> descs(3);
        const std::size_t nb = 20 * 24;
        std::vector<void*> dev(3);
        for (int k = 0; k < 3; ++k) {
            vmx_problem p = vi::integrator_crespo2021(150 + 50 * k, 6 + k, 40 + k).base();
            p.scalar = VMX_F32;
            p.bin_dims = 2;
            p.res[0] = 20;
            p.res[1] = 24;
            p.lo[0] = p.lo[1] = 0.0;
            p.hi[0] = p.hi[1] = 1.0;
            probs[k] = p;
            descs[k] = fs[k].desc;
            hip_ok(hipMalloc(&dev[k], nb * sizeof(float)));
        }
        ctx.check(vmx_integrate_batch(ctx.get(), 3, probs.data(), descs.data(), dev.data(), nullptr, nullptr));
        hip_ok(hipDeviceSynchronize());
        int same = 1;
        for (int k = 0; k < 3; ++k) {
            std::vector<float> batched(nb), alone(nb, 0.f);
            hip_ok(hipMemcpy(batched.data(), dev[k], nb * sizeof(float), hipMemcpyDeviceToHost));
            ctx.check(vmx_integrate(ctx.get(), &probs[k], &descs[k], alone.data(), nullptr));
            same &= int(same_bits(batched, alone));
            hip_ok(hipFree(dev[k]));
        }
        std::printf("batch3_f32 %d\n", same);
    }

    // 6. training pairs of a user integrand: those of the equal built-in polynomial
    {
        uint64_t np = 0;
        int32_t nf = 0, kp = 0;
        vi::Context& ctx = vi::Context::thread_default();
        ctx.check(vmx_refiner_size(2, &np, &nf, &kp));
        vmx_problem p = vi::integrator_crespo2021(255, 0, 0).base();
        p.scalar = VMX_F32;
        p.bin_dims = 2;
        p.res[0] = p.res[1] = 32;
        p.lo[0] = p.lo[1] = 0.0;
        p.hi[0] = p.hi[1] = 1.0;
        vmx_train_opts o{};
        o.target = 0;
        o.stride = 1;
        const uint64_t cap = 4096;
        std::vector<uint16_t> xa(cap * kp), xb(cap * kp);
        std::vector<float> ta(cap), tb(cap);
        uint64_t na = 0, nb = 0;
        ctx.check(vmx_refiner_pairs(ctx.get(), 1, &p, &f2.desc, &o, cap, xa.data(), ta.data(), &na));
        ctx.check(vmx_refiner_pairs(ctx.get(), 1, &p, &b2.desc, &o, cap, xb.data(), tb.data(), &nb));
        // rows come in the order the rounds placed the nodes (atomic slots): compare the
        // sets of (features, target) rows
        auto rows = [&](const std::vector<uint16_t>& x, const std::vector<float>& tg, uint64_t n) {
            std::vector<std::string> r(n);
            for (uint64_t i = 0; i < n; ++i) {
                r[i].assign(reinterpret_cast<const char*>(&x[i * kp]), sizeof(uint16_t) * kp);
                r[i].append(reinterpret_cast<const char*>(&tg[i]), sizeof(float));
            }
            std::sort(r.begin(), r.end());
            return r;
        };
        std::printf("pairs_f32 %d %llu\n", int(na == nb && rows(xa, ta, na) == rows(xb, tb, nb)),
                    (unsigned long long)na);
    }

    // 7. the Fubini form (integrator_crespo2021_infinite<2> over a finite 3-D range) with a
    //    user functor as its inner integrand: the built-in polynomial's bins bit for bit
    {
        const PolyFn<3> p3{3, {1.0, 0.5, 0.25, 0.0}, {{1, 2, 0}, {2, 0, 1}, {0, 0, 0}, {0, 0, 0}}, {1.0, 1.0, 1.0},
                           1.5, 0.25};
        const auto it = vi::integrator_crespo2021_infinite<2>(300, 4, 8, 6);
        Grid<float> a(16, 12), b(16, 12);
        vi::integrate(it, a, std::array<std::size_t, 2>{16, 12}, vi::device_integrand_fubini<3, 2>(p3),
                      vi::range_primary<3>());
        vi::integrate(it, b, std::array<std::size_t, 2>{16, 12}, vi::poly(poly_params(p3), 3), vi::range_primary<3>());
        std::printf("fubini3_f32 %d\n", int(same_bits(a.v, b.v)));
    }

    // 8. the Fubini form over an infinite range with a functor of a sequence: equal to
    //    the built-in Dyson series bit for bit (crespo2021_infinite at 2 and 4 region dims,
    //    the optimised Fubini integrator's stratified form at 1), and the reference's own
    //    documented series against its closed form
    const auto dys = vi::device_integrand_infinite<1, 2, 4>(DysonFn{0.75f});
    {
        const auto it = vi::integrator_crespo2021_infinite<2>(300, 4, 8, 6);
        Grid<float> a(16, 12), b(16, 12);
        vi::integrate(it, a, std::array<std::size_t, 2>{16, 12}, dys, vi::range_primary_infinite<float>());
        vi::integrate(it, b, std::array<std::size_t, 2>{16, 12}, vi::dyson(0.75), vi::range_primary_infinite<float>());
        std::printf("dyson_seq2_f32 %d\n", int(same_bits(a.v, b.v)));
    }
    {
        const auto it = vi::integrator_adaptive_fubini_variance_reduction_parallel_optimized<1>(100, 4, 16, 9);
        const auto rg = vi::range_infinite<float>({0.0f, 0.25f}, {1.0f, 0.75f});
        std::vector<float> a(32, 0.f), b(32, 0.f);
        struct V {
            std::vector<float>* v;
            float& operator()(const std::array<std::size_t, 1>& i) const { return (*v)[i[0]]; }
        };
        V va{&a}, vb{&b};
        vi::integrate(it, va, std::array<std::size_t, 1>{32}, dys, rg);
        vi::integrate(it, vb, std::array<std::size_t, 1>{32}, vi::dyson(0.75), rg);
        std::printf("dyson_seq1_opt_f32 %d\n", int(same_bits(a, b)));
    }
    {
        // the reference's own program (main/compilation-tests/crespo21.cc:27-36): 4 region
        // dims, 10 bins, through the reference's integrate(integrator, bins, f, range)
        auto range_infinite = vi::range_primary_infinite<float>();
        std::vector<float> a(10, 0.0f), b(10, 0.0f);
        vi::integrate(vi::integrator_crespo2021_infinite<4>(1280, 8, 256), a, dys, range_infinite);
        vi::integrate(vi::integrator_crespo2021_infinite<4>(1280, 8, 256), b, vi::dyson(0.75), range_infinite);
        std::printf("dyson_seq4_f32 %d\n", int(same_bits(a, b)));
    }
    {
        const auto it = vi::integrator_crespo2021_infinite<1>(63, 16, 4096, 5);
        std::vector<float> a(16, 0.f);
        struct V {
            std::vector<float>* v;
            float& operator()(const std::array<std::size_t, 1>& i) const { return (*v)[i[0]]; }
        };
        V va{&a};
        vi::integrate(it, va, std::array<std::size_t, 1>{16}, vi::device_integrand_infinite<1>(SeriesFn{0.5f}),
                      vi::range_primary_infinite<float>());
        save(out, "series_seq1", a);
        std::printf("series_seq1 1\n");
    }
    // a sequence functor is refused by the hooks where the form's inner_dim is not an
    // infinite range, a D-dim functor where it is, and region dims not instantiated
    {
        auto msg_of = [&](const vi::Integrand& inner, int inner_dim) {
            vmx_fubini_params fp{};
            fp.inner_family = VMX_FAMILY_CUSTOM;
            fp.inner_dim = inner_dim;
            fp.mc_samples = 2;
            for (int d = 0; d < VMX_MAX_DIM; ++d) fp.rest_hi[d] = 1.0;
            fp.inner = inner.desc.params;
            vi::Integrand g;
            g.desc = vmx_integrand{VMX_FAMILY_FUBINI, 2, &fp};
            Grid<float> gr(4, 4);
            std::string msg = "none";
            try {
                vi::integrate(vi::integrator_crespo2021(15, 2), gr, std::array<std::size_t, 2>{4, 4}, g,
                              vi::range_primary<2>());
            } catch (const std::exception& e) {
                msg = e.what();
            }
            return msg;
        };
        const PolyFn<3> p3{1, {1.0, 0.0, 0.0, 0.0}, {{1, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, {1.0, 1.0, 1.0},
                           0.0, 0.0};
        const std::string m1 = msg_of(dys, 3);
        const std::string m2 = msg_of(vi::device_integrand_fubini<3, 2>(p3), VMX_INFINITE_DIM);
        std::string m3 = "none";
        try {
            Grid<float> gr(4, 4, 4);
            vi::integrate(vi::integrator_crespo2021_infinite<3>(15, 2, 2, 0), gr, std::array<std::size_t, 3>{4, 4, 4},
                          dys, vi::range_primary_infinite<float>());
        } catch (const std::exception& e) {
            m3 = e.what();
        }
        std::printf("seq_refused %d\n", int(m1.find("infinite range only") != std::string::npos &&
                                             m2.find("inner_dim differs") != std::string::npos &&
                                             m3.find("not instantiated") != std::string::npos));
    }

    // 9. bins of a type the functor was not instantiated for are refused
    {
        Grid<double> g(4, 4);
        std::string msg = "none";
        try {
            vi::integrate(vi::integrator_crespo2021(15, 2), g, std::array<std::size_t, 2>{4, 4}, ft,
                          vi::range_primary<2>());
        } catch (const std::exception& e) {
            msg = e.what();
        }
        std::printf("unsupported %d\n", int(msg.find("not instantiated for double") != std::string::npos));
    }
    return 0;
}
