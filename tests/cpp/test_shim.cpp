// Exercises the header-only boost::compute templates (include/boost/compute)
// exactly as reference-side user code would, on the reference's own golden
// vectors, and cross-checks large inputs against std:: algorithms.  Exit
// status = number of failed checks.  Built by tests/cpp/Makefile.
#include <boost/compute.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

namespace bc = boost::compute;

static int failures = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            ++failures;                                                           \
        }                                                                         \
    } while (0)

template <class T>
static std::vector<T> host(const bc::vector<T> &v, bc::command_queue &q)
{
    std::vector<T> h(v.size());
    bc::copy(v.begin(), v.end(), h.begin(), q);
    return h;
}

int main()
{
    bc::device dev = bc::system::default_device();
    bc::command_queue &queue = bc::system::default_queue();
    std::printf("device %d: %s (%s), %u CUs\n", dev.id(), dev.name().c_str(), dev.arch().c_str(), dev.compute_units());

    {  // sort: test/test_sort.cpp:135 vector
        int data[] = {-4, 152, -5000, 963, 75321, -456, 0, 1112};
        bc::vector<int> v(data, data + 8, queue);
        bc::sort(v.begin(), v.end(), queue);
        CHECK(host(v, queue) == (std::vector<int>{-5000, -456, -4, 0, 152, 963, 1112, 75321}));
        bc::sort(v.begin(), v.end(), bc::greater<int>(), queue);
        CHECK(host(v, queue) == (std::vector<int>{75321, 1112, 963, 152, 0, -4, -456, -5000}));
    }
    {  // detail::radix_sort on a sub-range: test/test_radix_sort.cpp:530
        int data[] = {9, 8, 7, 6, 5, 4, 3, 2, 1, 0};
        bc::vector<int> v(data, data + 10, queue);
        bc::detail::radix_sort(v.begin() + 2, v.end() - 2, queue);
        CHECK(host(v, queue) == (std::vector<int>{9, 8, 2, 3, 4, 5, 6, 7, 1, 0}));
    }
    {  // float radix order puts -0.0 before +0.0
        float data[] = {-6023.0f, 152.5f, -63.0f, 1234567.0f, 11.2f, -5000.1f, 0.0f, 14.0f, -8.25f, -0.0f};
        bc::vector<float> v(data, data + 10, queue);
        bc::detail::radix_sort(v.begin(), v.end(), queue);
        std::vector<float> h = host(v, queue);
        CHECK(std::is_sorted(h.begin(), h.end()));
        CHECK(std::signbit(h[4]) && !std::signbit(h[5]) && h[4] == 0.0f && h[5] == 0.0f);
    }
    {  // radix_sort_by_key is stable: test/test_radix_sort_by_key.cpp:40
        int k[] = {10, 9, 2, 7, 6, -1, 4, 2, 2, 10};
        int v[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
        bc::vector<int> keys(k, k + 10, queue), vals(v, v + 10, queue);
        bc::detail::radix_sort_by_key(keys.begin(), keys.end(), vals.begin(), queue);
        CHECK(host(keys, queue) == (std::vector<int>{-1, 2, 2, 2, 4, 6, 7, 9, 10, 10}));
        CHECK(host(vals, queue) == (std::vector<int>{6, 3, 8, 9, 7, 5, 4, 2, 1, 10}));
    }
    {  // sort_by_key, 1024 reversed keys: test/test_sort_by_key.cpp:82
        const int n = 1024;
        std::vector<int> hk(n);
        std::vector<float> hv(n);
        for (int i = 0; i < n; ++i) {
            hk[i] = n - i;
            hv[i] = (n - i) / 2.f;
        }
        bc::vector<int> keys(hk.begin(), hk.end(), queue);
        bc::vector<float> vals(hv.begin(), hv.end(), queue);
        bc::sort_by_key(keys.begin(), keys.end(), vals.begin(), queue);
        std::vector<int> gk = host(keys, queue);
        std::vector<float> gv = host(vals, queue);
        bool ok = true;
        for (int i = 0; i < n; ++i) ok = ok && gk[i] == i + 1 && gv[i] == (i + 1) / 2.f;
        CHECK(ok);
    }
    {  // stable_sort_by_key keeps tie order: test/test_stable_sort_by_key.cpp:165
        const int n = 128;
        std::vector<int> k(n), v(n);
        for (int i = 0; i < n; ++i) k[i] = v[i] = -i;
        k[n / 2] = k[n - 2] = k[n - 1] = -256;
        v[n / 2] = 3;
        v[n - 2] = 1;
        v[n - 1] = 2;
        bc::vector<int> keys(k.begin(), k.end(), queue), vals(v.begin(), v.end(), queue);
        bc::stable_sort_by_key(keys.begin(), keys.end(), vals.begin(), queue);
        std::vector<int> gv = host(vals, queue);
        CHECK(gv[0] == 3 && gv[1] == 1 && gv[2] == 2);
    }
    {  // 1M uint32 keys against std::sort, and a host-iterator sort
        std::mt19937 g(0xB0057001u);
        std::vector<unsigned> h(1 << 20);
        for (auto &x : h) x = g();
        bc::vector<unsigned> v(h.begin(), h.end(), queue);
        bc::sort(v.begin(), v.end(), queue);
        std::vector<unsigned> ref = h;
        std::sort(ref.begin(), ref.end());
        CHECK(host(v, queue) == ref);
        std::vector<unsigned> hh = h;
        bc::sort(hh.begin(), hh.end(), queue);
        CHECK(hh == ref);
        bc::vector<unsigned> s(h.begin(), h.end(), queue);
        bc::stable_sort(s.begin(), s.end(), bc::greater<unsigned>(), queue);
        std::sort(ref.begin(), ref.end(), std::greater<unsigned>());
        CHECK(host(s, queue) == ref);
    }
    {  // scans: test/test_scan.cpp:41, :93, :329, :359; test/test_partial_sum.cpp:28
        int data[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
        bc::vector<int> v(data, data + 12, queue), r(12, queue.get_context());
        bc::inclusive_scan(v.begin(), v.end(), r.begin(), queue);
        CHECK(host(r, queue) == (std::vector<int>{0, 1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 66}));
        bc::exclusive_scan(v.begin(), v.end(), v.begin(), queue);  // in place
        CHECK(host(v, queue) == (std::vector<int>{0, 0, 1, 3, 6, 10, 15, 21, 28, 36, 45, 55}));
        int m[] = {1, 2, 1, 2, 3};
        bc::vector<int> mv(m, m + 5, queue), mo(5, queue.get_context());
        bc::inclusive_scan(mv.begin(), mv.end(), mo.begin(), bc::multiplies<int>(), queue);
        CHECK(host(mo, queue) == (std::vector<int>{1, 2, 2, 4, 12}));
        bc::exclusive_scan(mv.begin(), mv.end(), mo.begin(), int(10), bc::multiplies<int>(), queue);
        CHECK(host(mo, queue) == (std::vector<int>{10, 10, 20, 20, 40}));
        int p[] = {1, 2, 5, 3, 9, 1, 4, 2};
        bc::vector<int> pa(p, p + 8, queue), pb(8, queue.get_context());
        bc::vector<int>::iterator end = bc::partial_sum(pa.begin(), pa.end(), pb.begin(), queue);
        CHECK(end == pb.end());
        CHECK(host(pb, queue) == (std::vector<int>{1, 3, 8, 11, 20, 21, 25, 27}));
    }
    {  // reduce / accumulate: test/test_reduce.cpp:29, :77, :88; test/test_accumulate.cpp:25, :42
        int data[] = {1, 5, 9, 13, 17};
        bc::vector<int> v(data, data + 5, queue);
        int sum = 0, prod = 0;
        bc::reduce(v.begin(), v.end(), &sum, bc::plus<int>(), queue);
        bc::reduce(v.begin(), v.end(), &prod, bc::multiplies<int>(), queue);
        CHECK(sum == 45 && prod == 9945);
        int d8[] = {1, 2, 3, 4, 5, 6, 7, 8};
        bc::vector<int> in(d8, d8 + 8, queue), res(2, queue.get_context());
        bc::reduce(in.begin(), in.begin() + 4, res.begin(), queue);
        bc::reduce(in.begin() + 4, in.end(), res.end() - 1, queue);
        CHECK(host(res, queue) == (std::vector<int>{10, 26}));
        int mm[] = {11, 5, 92, 13, 42};
        bc::vector<int> mv(mm, mm + 5, queue);
        int mn = 0, mx = 0;
        bc::reduce(mv.begin(), mv.end(), &mn, bc::min<int>(), queue);
        bc::reduce(mv.begin(), mv.end(), &mx, bc::max<int>(), queue);
        CHECK(mn == 5 && mx == 92);
        int a[] = {2, 4, 6, 8};
        bc::vector<int> av(a, a + 4, queue);
        CHECK(bc::accumulate(av.begin(), av.end(), 0, queue) == 20);
        CHECK(bc::accumulate(av.begin(), av.end(), -10, queue) == 10);
        CHECK(bc::accumulate(av.begin(), av.end(), 2, bc::multiplies<int>(), queue) == 768);
        std::vector<float> f(100000, 1.01f);
        bc::vector<float> fv(f.begin(), f.end(), queue);
        float fs = bc::accumulate(fv.begin(), fv.end(), 0.f, queue);
        CHECK(std::fabs(fs - 101000.0f) <= 0.01f);  // 1.01f * 1e5 in exact arithmetic, rounded
    }
    {  // reduce_by_key: test/test_reduce_by_key.cpp:26 and the 1024-element runs (:48-73)
        int k[] = {0, 2, -3, -3, -3, -3, -3, 4};
        int d[] = {1, 1, 1, 1, 1, 2, 5, 1};
        bc::vector<int> ki(k, k + 8, queue), vi(d, d + 8, queue), ko(8, queue.get_context()), vo(8, queue.get_context());
        auto ends = bc::reduce_by_key(ki.begin(), ki.end(), vi.begin(), ko.begin(), vo.begin(), queue);
        CHECK(ends.first == ko.begin() + 4 && ends.second == vo.begin() + 4);
        std::vector<int> gk = host(ko, queue), gv = host(vo, queue);
        CHECK(std::vector<int>(gk.begin(), gk.begin() + 4) == (std::vector<int>{0, 2, -3, 4}));
        CHECK(std::vector<int>(gv.begin(), gv.begin() + 4) == (std::vector<int>{1, 1, 10, 1}));
        bc::vector<int> lk(1024, 0, queue), lv(1024, 1, queue), lko(1024, queue.get_context()), lvo(1024, queue.get_context());
        lk[137] = 1;
        lk[677] = 1;
        lk[1001] = 1;
        bc::inclusive_scan(lk.begin(), lk.end(), lk.begin(), queue);
        auto e2 = bc::reduce_by_key(lk.begin(), lk.end(), lv.begin(), lko.begin(), lvo.begin(), queue);
        CHECK(e2.first - lko.begin() == 4);
        std::vector<int> rv = host(lvo, queue);
        CHECK(rv[0] == 137 && rv[1] == 540 && rv[2] == 324 && rv[3] == 23);
    }
    {  // transform: test/test_transform.cpp:75; saxpy lambda a*_1+_2
        int a[] = {1, 2, 3, 4}, b[] = {10, 20, 30, 40};
        bc::vector<int> va(a, a + 4, queue), vb(b, b + 4, queue), out(4, queue.get_context());
        bc::transform(va.begin(), va.end(), vb.begin(), out.begin(), bc::plus<int>(), queue);
        CHECK(host(out, queue) == (std::vector<int>{11, 22, 33, 44}));
        bc::transform(va.begin(), va.end(), vb.begin(), out.begin(), bc::multiplies<int>(), queue);
        CHECK(host(out, queue) == (std::vector<int>{10, 40, 90, 160}));
        using bc::lambda::_1;
        using bc::lambda::_2;
        std::vector<float> x(1000), y(1000);
        for (int i = 0; i < 1000; ++i) {
            x[i] = 0.25f * i;
            y[i] = 1.0f - 0.5f * i;
        }
        bc::vector<float> dx(x.begin(), x.end(), queue), dy(y.begin(), y.end(), queue), dr(1000, queue.get_context());
        bc::transform(dx.begin(), dx.end(), dy.begin(), dr.begin(), 2.5f * _1 + _2, queue);
        std::vector<float> r = host(dr, queue);
        bool ok = true;
        for (int i = 0; i < 1000; ++i) ok = ok && r[i] == std::fma(2.5f, x[i], y[i]);
        CHECK(ok);
    }
    {  // errors surface as opencl_error with the OpenCL-valued code
        bool thrown = false;
        try {
            bc::command_queue bad(bc::device(1000));
        } catch (const bc::opencl_error &e) {
            thrown = e.error_code() == -1;
        }
        CHECK(thrown);
    }
    queue.finish();
    std::printf("%d failures\n", failures);
    return failures;
}
