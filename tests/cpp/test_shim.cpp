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
#include <string>
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
    {  // copy_if / copy_index_if: test/test_copy_if.cpp:24-67,101-117
        using bc::_1;
        int data[] = {1, 6, 3, 5, 8, 2, 4};
        bc::vector<int> input(data, data + 7, queue);
        bc::vector<int> output(7, queue.get_context());
        bc::fill(output.begin(), output.end(), -1, queue);
        bc::vector<int>::iterator it = bc::copy_if(input.begin(), input.end(), output.begin(), _1 < 5, queue);
        CHECK(it == output.begin() + 4);
        CHECK(host(output, queue) == (std::vector<int>{1, 3, 2, 4, -1, -1, -1}));
        bc::fill(output.begin(), output.end(), 42, queue);
        it = bc::copy_if(input.begin(), input.end(), output.begin(), _1 * 2 >= 10, queue);
        CHECK(it == output.begin() + 3);
        CHECK(host(output, queue) == (std::vector<int>{6, 5, 8, 42, 42, 42, 42}));
        bc::fill(output.begin(), output.end(), -1, queue);
        it = bc::detail::copy_index_if(input.begin(), input.end(), output.begin(), _1 < 5, queue);
        CHECK(it == output.begin() + 4);
        CHECK(host(output, queue) == (std::vector<int>{0, 2, 5, 6, -1, -1, -1}));
        int odd[] = {1, 2, 3, 4, 5, 1, 2, 3, 4, 5};
        bc::vector<int> in2(odd, odd + 10, queue), odds(10, queue.get_context()), evens(10, queue.get_context());
        CHECK(bc::copy_if(in2.begin(), in2.end(), odds.begin(), _1 % 2 == 1, queue) == odds.begin() + 6);
        CHECK(bc::copy_if(in2.begin(), in2.end(), evens.begin(), _1 % 2 == 0, queue) == evens.begin() + 4);
        std::vector<int> ho = host(odds, queue), he = host(evens, queue);
        CHECK(std::vector<int>(ho.begin(), ho.begin() + 6) == (std::vector<int>{1, 3, 5, 1, 3, 5}));
        CHECK(std::vector<int>(he.begin(), he.begin() + 4) == (std::vector<int>{2, 4, 2, 4}));
    }
    {  // unique / unique_copy: test/test_unique.cpp:26-51, test/test_unique_copy.cpp:23-34
        int data[] = {1, 6, 6, 4, 2, 2, 4};
        bc::vector<int> input(data, data + 7, queue);
        bc::vector<int> result(5, queue.get_context());
        CHECK(bc::unique_copy(input.begin(), input.end(), result.begin(), queue) == result.begin() + 5);
        CHECK(host(result, queue) == (std::vector<int>{1, 6, 4, 2, 4}));
        CHECK(bc::unique(input.begin(), input.end(), queue) == input.begin() + 5);
        CHECK(host(input, queue) == (std::vector<int>{1, 6, 4, 2, 4, 2, 4}));
        bc::vector<float> same(1024, queue.get_context());
        bc::fill(same.begin(), same.end(), 3.14f, queue);
        CHECK(bc::unique(same.begin(), same.end(), bc::equal_to<float>(), queue) == same.begin() + 1);
        CHECK(same.begin().read(queue) == 3.14f);
    }
    {  // partition / stable_partition / partition_copy: test/test_partition.cpp:27-80,
       // test/test_stable_partition.cpp:25-36
        float fdata[] = {1.0f, 2.0f, -1.0f, -2.0f, 3.0f, 4.0f, -3.0f, -4.0f};
        bc::vector<float> v(fdata, fdata + 8, queue);
        bc::vector<float>::iterator it = bc::partition(v.begin(), v.end(), bc::signbit_<float>(), queue);
        CHECK(it == v.begin() + 4);
        std::vector<float> h = host(v, queue);
        CHECK(h[0] < 0 && h[1] < 0 && h[2] < 0 && h[3] < 0 && h[4] > 0 && h[5] > 0 && h[6] > 0 && h[7] > 0);
        int idata[] = {1, 1, -2, 0, 5, -1, 2, 4, 0, -1};
        bc::vector<int> vi(idata, idata + 10, queue);
        bc::vector<int>::iterator iit = bc::stable_partition(vi.begin(), vi.begin() + 10, bc::_1 > 0, queue);
        CHECK(iit == vi.begin() + 5);
        CHECK(host(vi, queue) == (std::vector<int>{1, 1, 5, 2, 4, -2, 0, -1, 0, -1}));
        bc::vector<int> src(idata, idata + 10, queue), t(10, queue.get_context()), f(10, queue.get_context());
        std::pair<bc::vector<int>::iterator, bc::vector<int>::iterator> ends =
            bc::partition_copy(src.begin(), src.end(), t.begin(), f.begin(), bc::_1 > 0, queue);
        CHECK(ends.first == t.begin() + 5 && ends.second == f.begin() + 5);
        std::vector<int> ht = host(t, queue), hf = host(f, queue);
        CHECK(std::vector<int>(ht.begin(), ht.begin() + 5) == (std::vector<int>{1, 1, 5, 2, 4}));
        CHECK(std::vector<int>(hf.begin(), hf.begin() + 5) == (std::vector<int>{-2, 0, -1, 0, -1}));
    }
    {  // remove / remove_if: test/test_remove.cpp:23-53
        int data[] = {1, 2, 1, 3, 2, 4, 3, 4, 5};
        bc::vector<int> v(data, data + 9, queue);
        bc::vector<int>::iterator it = bc::remove(v.begin(), v.end(), 2, queue);
        CHECK(it == v.begin() + 7);
        std::vector<int> h = host(v, queue);
        CHECK(std::vector<int>(h.begin(), h.begin() + 7) == (std::vector<int>{1, 1, 3, 4, 3, 4, 5}));
        it = bc::remove(v.begin(), v.begin() + 7, 4, queue);
        CHECK(it == v.begin() + 5);
        it = bc::remove_if(v.begin(), v.begin() + 5, bc::_1 == 1, queue);
        CHECK(it == v.begin() + 3);
        h = host(v, queue);
        CHECK(std::vector<int>(h.begin(), h.begin() + 3) == (std::vector<int>{3, 3, 5}));
    }
    {  // count / count_if: test/test_count.cpp:32-42,63-72,117-157,206-218
        int data[] = {1, 2, 1, 2, 3};
        bc::vector<int> v(data, data + 5, queue);
        CHECK(bc::count(v.begin(), v.end(), 1, queue) == 2);
        CHECK(bc::count(v.begin(), v.end(), 3, queue) == 1);
        CHECK(bc::count(v.begin() + 1, v.end(), 1, queue) == 1);
        CHECK(bc::count(v.begin() + 1, v.end() - 1, 2, queue) == 2);
        float fdata[] = {1.0f, 2.5f, -1.0f, 3.0f, 5.0f, -8.0f};
        bc::vector<float> fv(fdata, fdata + 5, queue);
        CHECK(bc::count_if(fv.begin(), fv.end(), bc::_1 > 2.0f, queue) == 3);
        std::string s = "abcdefg\nhijklmn\nopqrs\ntuv\nwxyz\n";
        bc::vector<char> cv(s.begin(), s.end(), queue);
        CHECK(bc::count(cv.begin(), cv.end(), '\n', queue) == 5);
        unsigned char ud[] = {0x00, 0x10, 0x2F, 0x10, 0x01, 0x00, 0x01, 0x00};
        bc::vector<unsigned char> uv(ud, ud + 8, queue);
        CHECK(bc::count(uv.begin(), uv.end(), 0x00, queue) == 3);
        CHECK(bc::count(uv.begin(), uv.end(), 0x2F, queue) == 1);
        CHECK(bc::count(uv.begin(), uv.end(), 0xFF, queue) == 0);
        std::vector<int> iota(2048);
        std::iota(iota.begin(), iota.end(), 0);
        bc::vector<int> iv(iota.begin(), iota.end(), queue);
        CHECK(bc::count_if(iv.begin(), iv.end(), bc::_1 > 1024, queue) == 1023);
        CHECK(bc::count_if(iv.begin(), iv.end(), bc::_1 & 1, queue) == 1024);  // nonzero is true, as in C
        CHECK(bc::count_if(iv.begin(), iv.end(), (bc::_1 & 1) != 0, queue) == 1024);
        CHECK(bc::count_if(iv.begin(), iv.end(), bc::not1(bc::_1 > 1024), queue) == 1025);
    }
    {  // transform_reduce / inner_product: test/test_transform_reduce.cpp:23-99,
       // test/test_inner_product.cpp:23-52
        int data[] = {1, -2, -3, -4, 5};
        bc::vector<int> vec(data, data + 5, queue);
        int sum = 0;
        bc::transform_reduce(vec.begin(), vec.end(), &sum, bc::abs<int>(), bc::plus<int>(), queue);
        CHECK(sum == 15);
        float fdata[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
        bc::vector<float> fv(fdata, fdata + 10, queue);
        float fsum = 0;
        bc::reduce(fv.begin(), fv.end(), &fsum, bc::plus<float>(), queue);
        const float mean = fsum / fv.size();
        CHECK(std::fabs(mean - 5.5f) < 1e-5f);
        using bc::lambda::_1;
        using bc::lambda::pow;
        bc::transform_reduce(fv.begin(), fv.end(), &fsum, pow(_1 - mean, 2), bc::plus<float>(), queue);
        CHECK(std::fabs(fsum / fv.size() - 8.25f) < 1e-4f);
        int d1[] = {1, 2, 3, 4}, d2[] = {10, 20, 30, 40};
        bc::vector<int> in1(d1, d1 + 4, queue), in2(d2, d2 + 4, queue);
        CHECK(bc::inner_product(in1.begin(), in1.end(), in2.begin(), 0, queue) == 300);
        CHECK(bc::inner_product(bc::make_counting_iterator<int>(0), bc::make_counting_iterator<int>(100),
                                bc::make_counting_iterator<int>(0), 0, queue) == 328350);
        // the general form: init accumulated with transform(a, b)
        CHECK(bc::inner_product(in1.begin(), in1.end(), in2.begin(), 5, bc::plus<int>(), bc::minus<int>(), queue) ==
              5 + (1 - 10) + (2 - 20) + (3 - 30) + (4 - 40));
        bc::vector<int> res(1, queue.get_context());
        bc::transform_reduce(in1.begin(), in1.end(), in2.begin(), res.begin(), bc::multiplies<int>(), bc::max<int>(),
                             queue);
        CHECK(res.begin().read(queue) == 160);
    }
    {  // experimental::sort_by_transform: test/test_sort_by_transform.cpp:25-37, :75-108
        int data[] = {1, -2, 4, -3, 0, 5, -8, -9};
        bc::vector<int> v(data, data + 8, queue);
        bc::experimental::sort_by_transform(v.begin(), v.end(), bc::abs<int>(), bc::less<int>(), queue);
        CHECK(host(v, queue) == (std::vector<int>{0, 1, -2, -3, 4, 5, -8, -9}));
        float f[] = {1, 2, 3, 0, 9, 8, 7, 0, 4, 5, 6, 0, 0, 0, 0, 0};
        bc::vector<bc::float4_> fv(reinterpret_cast<bc::float4_ *>(f), reinterpret_cast<bc::float4_ *>(f) + 4,
                                   queue);
        bc::experimental::sort_by_transform(fv.begin(), fv.end(), bc::get<1>(), bc::less<float>(), queue);
        std::vector<bc::float4_> h = host(fv, queue);
        CHECK(h[0] == bc::float4_(0.f, 0.f, 0.f, 0.f) && h[1] == bc::float4_(1.f, 2.f, 3.f, 0.f) &&
              h[2] == bc::float4_(4.f, 5.f, 6.f, 0.f) && h[3] == bc::float4_(9.f, 8.f, 7.f, 0.f));
        // stable by the first component: the order test/test_merge_sort_gpu.cpp:362-369 expects
        bc::int2_ d2[] = {bc::int2_(8, 3), bc::int2_(5, 1), bc::int2_(2, 1), bc::int2_(6, 1),
                          bc::int2_(8, 1), bc::int2_(7, 1), bc::int2_(4, 1), bc::int2_(8, 2)};
        bc::vector<bc::int2_> iv(d2, d2 + 8, queue);
        bc::experimental::sort_by_transform(iv.begin(), iv.end(), bc::get<0>(), bc::less<int>(), queue);
        std::vector<bc::int2_> hi = host(iv, queue);
        CHECK(hi[0] == bc::int2_(2, 1) && hi[1] == bc::int2_(4, 1) && hi[2] == bc::int2_(5, 1) &&
              hi[3] == bc::int2_(6, 1) && hi[4] == bc::int2_(7, 1) && hi[5] == bc::int2_(8, 3) &&
              hi[6] == bc::int2_(8, 1) && hi[7] == bc::int2_(8, 2));
    }
    {  // copy_async through page-locked host memory, with the queue busy meanwhile
        const size_t n = 1 << 22;
        unsigned *pin = 0, *back = 0;
        CHECK(bcmi_host_malloc(n * sizeof(unsigned), (void **)&pin) == 0);
        CHECK(bcmi_host_malloc(n * sizeof(unsigned), (void **)&back) == 0);
        for (size_t i = 0; i < n; ++i) pin[i] = (unsigned)(i * 2654435761u);
        bc::vector<unsigned> dv(n, queue.get_context());
        bc::future<bc::vector<unsigned>::iterator> f = bc::copy_async(pin, pin + n, dv.begin(), queue);
        CHECK(f.valid());
        CHECK(f.get() == dv.end());
        bc::sort(dv.begin(), dv.end(), queue);  // ordered after the copy on the queue
        bc::future<unsigned *> g = bc::copy_async(dv.begin(), dv.end(), back, queue);
        CHECK(g.get() == back + n);
        std::vector<unsigned> ref(pin, pin + n);
        std::sort(ref.begin(), ref.end());
        CHECK(std::equal(ref.begin(), ref.end(), back));
        CHECK(bcmi_host_free(pin) == 0 && bcmi_host_free(back) == 0);
    }
    {  // comparison sort, less/greater given to merge_sort_on_gpu: test/test_merge_sort_gpu.cpp:36-62
        using bc::char_;
        char_ data[] = {'c', 'a', '0', '7', 'B', 'F', '\0', '$'};
        bc::vector<char_> v(data, data + 8, queue);
        bc::detail::merge_sort_on_gpu(v.begin(), v.end(), bc::less<char_>(), queue);
        CHECK(host(v, queue) == (std::vector<char_>{'\0', '$', '0', '7', 'B', 'F', 'a', 'c'}));
        bc::detail::merge_sort_on_gpu(v.begin(), v.end(), bc::greater<char_>(), queue);
        CHECK(host(v, queue) == (std::vector<char_>{'c', 'a', 'F', 'B', '7', '0', '$', '\0'}));
    }
    {  // abs(a) < abs(b) and abs(a.x + a.y) < abs(b.x + b.y): test/test_merge_sort_gpu.cpp:224-296
        using bc::lambda::_1;
        using bc::lambda::_2;
        using bc::lambda::get;
        std::vector<int> d(1024);
        std::vector<bc::int2_> d2(1024);
        for (int i = 0; i < 1024; ++i) {
            d[i] = i % 2 ? 1024 - i : i - 1024;
            d2[i] = i % 2 ? bc::int2_(i, i) : bc::int2_(i - 1024, i - 1024);
        }
        bc::vector<int> v(d.begin(), d.end(), queue);
        bc::detail::merge_sort_on_gpu(v.begin(), v.end(), abs(_1) < abs(_2), queue);
        std::vector<int> h = host(v, queue);
        bool ok = true;
        for (size_t i = 1; i < h.size(); ++i) ok = ok && std::abs(h[i - 1]) <= std::abs(h[i]);
        CHECK(ok);
        bc::vector<bc::int2_> v2(d2.begin(), d2.end(), queue);
        bc::detail::merge_sort_on_gpu(v2.begin(), v2.end(),
                                      abs(get<0>(_1) + get<1>(_1)) < abs(get<0>(_2) + get<1>(_2)), queue);
        std::vector<bc::int2_> h2 = host(v2, queue);
        ok = true;
        for (size_t i = 1; i < h2.size(); ++i)
            ok = ok && std::abs(h2[i - 1][0] + h2[i - 1][1]) <= std::abs(h2[i][0] + h2[i][1]);
        CHECK(ok);
    }
    {  // a.s0 < b.s3 on long8_: test/test_merge_sort_gpu.cpp:299-330
        using bc::lambda::_1;
        using bc::lambda::_2;
        using bc::lambda::get;
        std::vector<bc::long8_> d(256);
        for (long i = 0; i < 256; ++i) d[i] = bc::long8_(i % 2 ? i : i * i);
        bc::vector<bc::long8_> v(d.begin(), d.end(), queue);
        bc::detail::merge_sort_on_gpu(v.begin(), v.end(), get<0>(_1) < get<3>(_2), queue);
        std::vector<bc::long8_> h = host(v, queue);
        bool ok = true;
        for (size_t i = 1; i < h.size(); ++i) ok = ok && h[i - 1][0] <= h[i][3];
        CHECK(ok);
    }
    {  // stable merge sort of int2_ by x: test/test_merge_sort_gpu.cpp:333-370
        using bc::lambda::_1;
        using bc::lambda::_2;
        using bc::lambda::get;
        bc::int2_ data[] = {bc::int2_(8, 3), bc::int2_(5, 1), bc::int2_(2, 1), bc::int2_(6, 1),
                            bc::int2_(8, 1), bc::int2_(7, 1), bc::int2_(4, 1), bc::int2_(8, 2)};
        bc::vector<bc::int2_> v(data, data + 8, queue);
        bc::detail::merge_sort_on_gpu(v.begin(), v.end(), get<0>(_1) < get<0>(_2), true, queue);
        std::vector<bc::int2_> h = host(v, queue);
        int want[8][2] = {{2, 1}, {4, 1}, {5, 1}, {6, 1}, {7, 1}, {8, 3}, {8, 1}, {8, 2}};
        bool ok = true;
        for (int i = 0; i < 8; ++i) ok = ok && h[i][0] == want[i][0] && h[i][1] == want[i][1];
        CHECK(ok);
    }
    {  // stable_sort of int2_ by each component: test/test_stable_sort.cpp:40-95
        using bc::lambda::_1;
        using bc::lambda::_2;
        using bc::lambda::get;
        bc::int2_ data[] = {bc::int2_(2, 1), bc::int2_(2, 2), bc::int2_(1, 2), bc::int2_(1, 1)};
        bc::vector<bc::int2_> w(data, data + 4, queue);
        bc::stable_sort(w.begin(), w.end(), get<0>(_1) < get<0>(_2), queue);
        std::vector<bc::int2_> h = host(w, queue);
        CHECK(h[0] == bc::int2_(1, 2) && h[1] == bc::int2_(1, 1) && h[2] == bc::int2_(2, 1) && h[3] == bc::int2_(2, 2));
        bc::stable_sort(w.begin(), w.end(), get<1>(_1) < get<1>(_2), queue);
        h = host(w, queue);
        CHECK(h[0] == bc::int2_(1, 1) && h[1] == bc::int2_(2, 1) && h[2] == bc::int2_(1, 2) && h[3] == bc::int2_(2, 2));
    }
    {  // sort by length(_1) < length(_2): test/test_sort.cpp:238-285
        using bc::lambda::_1;
        using bc::lambda::_2;
        float data[] = {1.0f, 0.2f, 1.3f, 1.0f, 6.7f, 0.0f, 5.2f, 3.4f, 1.4f, 1.4f};
        bc::vector<bc::float2_> v(reinterpret_cast<bc::float2_ *>(data), reinterpret_cast<bc::float2_ *>(data) + 5,
                                  queue);
        bc::sort(v.begin(), v.end(), length(_1) < length(_2), queue);
        bc::copy(v.begin(), v.end(), reinterpret_cast<bc::float2_ *>(data), queue);
        float want[] = {1.0f, 0.2f, 1.3f, 1.0f, 1.4f, 1.4f, 5.2f, 3.4f, 6.7f, 0.0f};
        CHECK(std::equal(data, data + 10, want));
    }
    {  // sort of a struct by a member: test/test_sort.cpp:293-326 (sort_by_x)
        struct Particle {
            float x, y;
        };
        std::vector<Particle> p = {{0.1f, 0.f}, {-0.4f, 0.f}, {10.0f, 0.f}, {0.001f, 0.f}};
        bc::vector<Particle> v(p.begin(), p.end(), queue);
        bc::sort(v.begin(), v.end(), bc::experimental::less_by(&Particle::x), queue);
        std::vector<Particle> h = host(v, queue);
        CHECK(h[0].x == -0.4f && h[1].x == 0.001f && h[2].x == 0.1f && h[3].x == 10.0f);
        bc::sort(v.begin(), v.end(), bc::experimental::greater_by(&Particle::x), queue);
        h = host(v, queue);
        CHECK(h[0].x == 10.0f && h[3].x == -0.4f);
    }
    {  // sort_by_key with a user comparator: test/test_sort_by_key.cpp:113-142
        using bc::lambda::_1;
        using bc::lambda::_2;
        const int n = 1024;
        std::vector<int> hk(n);
        std::vector<float> hv(n);
        for (int i = 0; i < n; ++i) {
            hk[i] = n - i;
            hv[i] = (n - i) / 2.f;
        }
        bc::vector<int> k(hk.begin(), hk.end(), queue);
        bc::vector<float> v(hv.begin(), hv.end(), queue);
        bc::sort_by_key(k.begin(), k.end(), v.begin(), _1 < _2, queue);
        std::vector<int> gk = host(k, queue);
        std::vector<float> gv = host(v, queue);
        CHECK(std::is_sorted(gk.begin(), gk.end()) && std::is_sorted(gv.begin(), gv.end()) && gk[0] == 1);
    }
    {  // stable_sort_by_key with a user comparator: test/test_stable_sort_by_key.cpp:165-208
        using bc::lambda::_1;
        using bc::lambda::_2;
        std::vector<int> hk(128), hv(128);
        for (int i = 0; i < 128; ++i) hk[i] = hv[i] = -i;
        hk[64] = hk[126] = hk[127] = -256;
        hv[64] = 3;
        hv[126] = 1;
        hv[127] = 2;
        bc::vector<int> k(hk.begin(), hk.end(), queue);
        bc::vector<int> v(hv.begin(), hv.end(), queue);
        bc::stable_sort_by_key(k.begin(), k.end(), v.begin(), _1 < _2, queue);
        std::vector<int> gk = host(k, queue), gv = host(v, queue);
        CHECK(gk[0] == -256 && gk[1] == -256 && gk[2] == -256 && gv[0] == 3 && gv[1] == 1 && gv[2] == 2);
        CHECK(std::is_sorted(gk.begin(), gk.end()));
    }
    {  // a comparator sort of a host vector (the reference maps it, sort.hpp:124-148), and
       // sort_by_key with _1 > _2 on a sub-range
        using bc::lambda::_1;
        using bc::lambda::_2;
        using bc::lambda::get;
        std::vector<bc::int2_> h = {bc::int2_(3, 0), bc::int2_(-1, 1), bc::int2_(3, 2), bc::int2_(0, 3),
                                    bc::int2_(-1, 4)};
        bc::sort(h.begin(), h.end(), get<0>(_1) < get<0>(_2), queue);
        CHECK(h[0] == bc::int2_(-1, 1) && h[1] == bc::int2_(-1, 4) && h[2] == bc::int2_(0, 3) &&
              h[3] == bc::int2_(3, 0) && h[4] == bc::int2_(3, 2));
        std::vector<int> hk = {9, 5, 7, 5, 1, 8, 2}, hv = {0, 1, 2, 3, 4, 5, 6};
        bc::vector<int> k(hk.begin(), hk.end(), queue), v(hv.begin(), hv.end(), queue);
        bc::sort_by_key(k.begin() + 1, k.end() - 1, v.begin() + 1, _1 > _2, queue);
        CHECK(host(k, queue) == (std::vector<int>{9, 8, 7, 5, 5, 1, 2}));
        CHECK(host(v, queue) == (std::vector<int>{0, 5, 2, 1, 3, 4, 6}));
    }
    {  // a large comparator sort against std::stable_sort: 3 M int2_ by y, descending
        using bc::lambda::_1;
        using bc::lambda::_2;
        using bc::lambda::get;
        const size_t n = 3000017;
        std::mt19937 gen(17);
        std::vector<bc::int2_> d(n);
        for (size_t i = 0; i < n; ++i) d[i] = bc::int2_((int)i, (int)(gen() % 20000) - 10000);
        bc::vector<bc::int2_> v(d.begin(), d.end(), queue);
        bc::stable_sort(v.begin(), v.end(), get<1>(_1) > get<1>(_2), queue);
        std::stable_sort(d.begin(), d.end(), [](const bc::int2_ &a, const bc::int2_ &b) { return a[1] > b[1]; });
        std::vector<bc::int2_> h = host(v, queue);
        bool ok = true;
        for (size_t i = 0; i < n; ++i) ok = ok && h[i][0] == d[i][0] && h[i][1] == d[i][1];
        CHECK(ok);
    }
    {  // mapped_view over host arrays: test/test_mapped_view.cpp:24-66
        int data[] = {1, 2, 3, 4, 5, 6, 7, 8};
        bc::mapped_view<int> view(data, 8, queue.get_context());
        bc::fill(view.begin(), view.end(), 4, queue);
        view.map(CL_MAP_READ, queue);
        CHECK(std::all_of(data, data + 8, [](int x) { return x == 4; }));
        view.unmap(queue);
        int d2[] = {5, 2, 3, 1, 8, 7, 4, 9};
        bc::mapped_view<int> v2(d2, 8, queue.get_context());
        bc::sort(v2.begin(), v2.end(), queue);
        v2.map(CL_MAP_READ, queue);
        CHECK(d2[0] == 1 && d2[7] == 9);
        v2.unmap(queue);
        int d3[] = {5, 2, 3, 1, 8, 7, 4, 9};
        bc::mapped_view<int> v3(d3, 8, queue.get_context());
        int sum = 0;
        bc::reduce(v3.begin(), v3.end(), &sum, queue);
        CHECK(sum == 39);
        // host writes under a write mapping reach the device at unmap
        v3.map(queue);
        d3[0] = 100;
        v3.unmap(queue);
        bc::reduce(v3.begin(), v3.end(), &sum, queue);
        CHECK(sum == 134);
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
    {  // a kernel that gives up a bounded look-back wait: the next finish
       // raises CL_OUT_OF_RESOURCES, the one after succeeds, and the queue
       // keeps working (the probe's lowered spin limit is restored)
        queue.finish();
        CHECK(bcmi_debug_fault_probe(queue.get(), 64) == 0);
        int code = 0;
        std::string what;
        try {
            queue.finish();
        } catch (const bc::opencl_error &e) {
            code = e.error_code();
            what = e.what();
        }
        CHECK(code == -5 && what.find("spin limit") != std::string::npos);
        bool clean = true;
        try {
            queue.finish();
        } catch (const bc::opencl_error &) {
            clean = false;
        }
        CHECK(clean);
        std::vector<int> h(100000);
        for (size_t i = 0; i < h.size(); ++i) h[i] = (int)((i * 2654435761u) % 100003u) - 50000;
        bc::vector<int> d(h.begin(), h.end(), queue);
        bc::sort(d.begin(), d.end(), queue);
        std::vector<int> got(h.size());
        bc::copy(d.begin(), d.end(), got.begin(), queue);
        std::sort(h.begin(), h.end());
        CHECK(got == h);
    }
    queue.finish();
    std::printf("%d failures\n", failures);
    return failures;
}
