// The fancy iterators (counting_iterator, constant_iterator,
// transform_iterator, zip_iterator) and vector-type elements through the
// header-only boost::compute templates, on the reference's own test
// vectors: test/test_scan.cpp:140-286, test/test_transform_reduce.cpp:24-90,
// test/test_reduce.cpp:207-235, test/test_zip_iterator.cpp:157-221,
// test/test_constant_iterator.cpp:60-91; plus larger generated ranges
// checked against std:: loops.  Exit status = number of failed checks.
#include <boost/compute.hpp>

#include <cmath>
#include <cstdio>
#include <numeric>
#include <tuple>
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

static bool close(float a, float b, float pct) { return std::fabs(a - b) <= std::fabs(b) * pct / 100.0f; }

int main()
{
    bc::command_queue &queue = bc::system::default_queue();
    bc::context context = queue.get_context();
    using bc::_1;

    {  // test_scan.cpp:140-164 inclusive_scan_int2
        int data[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 0};
        bc::vector<bc::int2_> input(reinterpret_cast<bc::int2_ *>(data), reinterpret_cast<bc::int2_ *>(data) + 5,
                                    queue);
        bc::vector<bc::int2_> output(5, context);
        bc::inclusive_scan(input.begin(), input.end(), output.begin(), queue);
        std::vector<bc::int2_> h = host(output, queue);
        CHECK(h[0] == bc::int2_(1, 2) && h[1] == bc::int2_(4, 6) && h[2] == bc::int2_(9, 12) &&
              h[3] == bc::int2_(16, 20) && h[4] == bc::int2_(25, 20));
        // exclusive, with an init vector
        bc::exclusive_scan(input.begin(), input.end(), output.begin(), bc::int2_(10, 100), queue);
        h = host(output, queue);
        CHECK(h[0] == bc::int2_(10, 100) && h[1] == bc::int2_(11, 102) && h[4] == bc::int2_(26, 120));
    }
    {  // test_scan.cpp:190-222 inclusive_scan_counting_iterator
        bc::vector<int> result(10, context);
        bc::inclusive_scan(bc::make_counting_iterator(1), bc::make_counting_iterator(11), result.begin(), queue);
        CHECK(host(result, queue) == (std::vector<int>{1, 3, 6, 10, 15, 21, 28, 36, 45, 55}));
    }
    {  // test_scan.cpp:224-256 exclusive_scan_counting_iterator
        bc::vector<int> result(10, context);
        bc::exclusive_scan(bc::make_counting_iterator(1), bc::make_counting_iterator(11), result.begin(), queue);
        CHECK(host(result, queue) == (std::vector<int>{0, 1, 3, 6, 10, 15, 21, 28, 36, 45}));
    }
    {  // test_scan.cpp:258-285 inclusive_scan_transform_iterator
        float data[] = {1.0f, 2.0f, 3.0f, 4.0f, 5.0f};
        bc::vector<float> input(data, data + 5, queue);
        bc::vector<float> output(5, context);
        bc::inclusive_scan(input.begin(), input.end(), output.begin(), queue);
        std::vector<float> h = host(output, queue);
        CHECK(close(h[0], 1.0f, 1e-4f) && close(h[1], 3.0f, 1e-4f) && close(h[4], 15.0f, 1e-4f));
        bc::inclusive_scan(bc::make_transform_iterator(input.begin(), pown(_1, 2)),
                           bc::make_transform_iterator(input.end(), pown(_1, 2)), output.begin(), queue);
        h = host(output, queue);
        CHECK(close(h[0], 1.0f, 1e-4f) && close(h[1], 5.0f, 1e-4f) && close(h[2], 14.0f, 1e-4f) &&
              close(h[3], 30.0f, 1e-4f) && close(h[4], 55.0f, 1e-4f));
    }
    {  // test_transform_reduce.cpp:24-40 sum_abs_int_doctest
        int data[] = {1, -2, -3, -4, 5};
        bc::vector<int> vec(data, data + 5, queue);
        int sum = 0;
        bc::transform_reduce(vec.begin(), vec.end(), &sum, bc::abs<int>(), bc::plus<int>(), queue);
        CHECK(sum == 15);
    }
    {  // test_transform_reduce.cpp:42-60 multiply_vector_length
        float data[] = {2.0f, 0.0f, 0.0f, 0.0f, 0.0f, 3.0f, 0.0f, 0.0f, 0.0f, 0.0f, 4.0f, 0.0f};
        bc::vector<bc::float4_> vector(reinterpret_cast<bc::float4_ *>(data),
                                       reinterpret_cast<bc::float4_ *>(data) + 3, queue);
        float product = 0.0f;
        bc::transform_reduce(vector.begin(), vector.end(), &product, bc::length<bc::float4_>(),
                             bc::multiplies<float>(), queue);
        CHECK(close(product, 24.0f, 1e-4f));
    }
    {  // test_transform_reduce.cpp:62-90 mean_and_std_dev
        float data[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
        bc::vector<float> vector(data, data + 10, queue);
        float sum = 0.0f;
        bc::reduce(vector.begin(), vector.end(), &sum, bc::plus<float>(), queue);
        float mean = sum / vector.size();
        CHECK(close(mean, 5.5f, 1e-4f));
        bc::transform_reduce(vector.begin(), vector.end(), &sum, bc::lambda::pow(_1 - mean, 2), bc::plus<float>(),
                             queue);
        CHECK(close(sum / vector.size(), 8.25f, 1e-4f));
    }
    {  // test_reduce.cpp:207-235 reduce_transform_iterator
        int data[] = {1, 3, 5, 7, 9};
        bc::vector<int> vector(data, data + 5, queue);
        int sum;
        bc::reduce(bc::make_transform_iterator(vector.begin(), _1 + 1),
                   bc::make_transform_iterator(vector.end(), _1 + 1), &sum, queue);
        CHECK(sum == 30);
        bc::reduce(bc::make_transform_iterator(vector.begin(), _1 > 4),
                   bc::make_transform_iterator(vector.end(), _1 > 4), &sum, bc::plus<int>(), queue);
        CHECK(sum == 3);
        bc::reduce(bc::make_transform_iterator(vector.begin(), _1 * _1),
                   bc::make_transform_iterator(vector.end(), _1 * _1), &sum, queue);
        CHECK(sum == 165);
    }
    {  // test_zip_iterator.cpp:157-207 zip_iterator_get
        int data1[] = {0, 2, 4, 6, 8};
        int data2[] = {1, 3, 5, 7, 9};
        bc::vector<int> input1(data1, data1 + 5, queue);
        bc::vector<int> input2(data2, data2 + 5, queue);
        bc::vector<int> output(5, context);
        bc::transform(bc::make_zip_iterator(std::make_tuple(input1.begin())),
                      bc::make_zip_iterator(std::make_tuple(input1.end())), output.begin(), bc::get<0>(), queue);
        CHECK(host(output, queue) == (std::vector<int>{0, 2, 4, 6, 8}));
        bc::transform(bc::make_zip_iterator(std::make_tuple(input2.begin(), input1.begin())),
                      bc::make_zip_iterator(std::make_tuple(input2.end(), input1.end())), output.begin(),
                      bc::get<0>(), queue);
        CHECK(host(output, queue) == (std::vector<int>{1, 3, 5, 7, 9}));
        bc::transform(bc::make_zip_iterator(std::make_tuple(input1.begin(), input2.begin(), input1.begin())),
                      bc::make_zip_iterator(std::make_tuple(input1.end(), input2.end(), input1.end())),
                      output.begin(), bc::get<1>(), queue);
        CHECK(host(output, queue) == (std::vector<int>{1, 3, 5, 7, 9}));
    }
    {  // test_zip_iterator.cpp:209-221 zip_constant_iterator
        bc::vector<int> result(4, context);
        bc::transform(bc::make_zip_iterator(std::make_tuple(bc::make_constant_iterator(7))),
                      bc::make_zip_iterator(std::make_tuple(bc::make_constant_iterator(7, result.size()))),
                      result.begin(), bc::get<0>(), queue);
        CHECK(host(result, queue) == (std::vector<int>{7, 7, 7, 7}));
    }
    {  // test_constant_iterator.cpp:60-91 copy / fill_with_copy_doctest
        bc::vector<int> vector(10, context);
        bc::copy(bc::make_constant_iterator(42, 0), bc::make_constant_iterator(42, 10), vector.begin(), queue);
        CHECK(host(vector, queue) == std::vector<int>(10, 42));
    }
    {  // test_copy.cpp:299-331 check_copy_type: int -> short on the device
        int data[] = {1, 2, 3, 4, 5, 6, 7, 8};
        bc::vector<int> a(data, data + 8, queue);
        bc::vector<short> c(8, context);
        bc::copy(a.begin(), a.end(), c.begin(), queue);
        CHECK(host(c, queue) == (std::vector<short>{1, 2, 3, 4, 5, 6, 7, 8}));
        // float -> int truncates toward zero, int -> float rounds to nearest
        float fd[] = {1.75f, -1.75f, 2.5f, -0.5f, 1e9f};
        bc::vector<float> f(fd, fd + 5, queue);
        bc::vector<int> fi(5, context);
        bc::copy(f.begin(), f.end(), fi.begin(), queue);
        CHECK(host(fi, queue) == (std::vector<int>{1, -1, 2, 0, 1000000000}));
        int big[] = {16777217, -16777219, 3};
        bc::vector<int> b(big, big + 3, queue);
        bc::vector<float> bf(3, context);
        bc::copy(b.begin(), b.end(), bf.begin(), queue);
        CHECK(host(bf, queue) == (std::vector<float>{16777216.0f, -16777220.0f, 3.0f}));
    }
    {  // scans whose output type differs from the input type: computed in
       // the input type, converted on store (scan_on_gpu.hpp:196-205)
        int data[] = {1, 2, 3, 4, 5};
        bc::vector<int> input(data, data + 5, queue);
        bc::vector<float> out(5, context);
        bc::inclusive_scan(input.begin(), input.end(), out.begin(), queue);
        CHECK(host(out, queue) == (std::vector<float>{1, 3, 6, 10, 15}));
        bc::exclusive_scan(input.begin(), input.end(), out.begin(), 100, queue);
        CHECK(host(out, queue) == (std::vector<float>{100, 101, 103, 106, 110}));
        // int sums that wrap: the int result, then converted
        const int n = 100003;
        std::vector<int> x(n);
        for (int i = 0; i < n; ++i) x[i] = (int)(((unsigned)i * 2654435761u) >> 1);
        bc::vector<int> dx(x.begin(), x.end(), queue);
        bc::vector<double> dd(n, context);
        bc::inclusive_scan(dx.begin(), dx.end(), dd.begin(), queue);
        std::vector<double> hd = host(dd, queue);
        unsigned acc = 0;
        bool ok = true;
        for (int i = 0; i < n; ++i) {
            acc += (unsigned)x[i];
            ok = ok && hd[i] == (double)(int)acc;
        }
        CHECK(ok);
        // a counting range scanned into float elements
        bc::vector<float> cf(10, context);
        bc::inclusive_scan(bc::make_counting_iterator(1), bc::make_counting_iterator(11), cf.begin(), queue);
        CHECK(host(cf, queue) == (std::vector<float>{1, 3, 6, 10, 15, 21, 28, 36, 45, 55}));
    }
    {  // larger generated and mapped ranges against std:: loops
        const int n = 300007;
        bc::vector<long> out(n, context);
        bc::inclusive_scan(bc::make_counting_iterator(-5000L), bc::make_counting_iterator(-5000L + n), out.begin(),
                           queue);
        std::vector<long> h = host(out, queue), e(n);
        long acc = 0;
        for (int i = 0; i < n; ++i) e[i] = (acc += -5000L + i);
        CHECK(h == e);
        std::vector<int> x(n);
        for (int i = 0; i < n; ++i) x[i] = (int)(((long)i * 7919) % 1001) - 500;
        bc::vector<int> dx(x.begin(), x.end(), queue), dy(n, context);
        bc::exclusive_scan(bc::make_transform_iterator(dx.begin(), bc::abs<int>()),
                           bc::make_transform_iterator(dx.end(), bc::abs<int>()), dy.begin(), 3, bc::max<int>(),
                           queue);
        std::vector<int> hy = host(dy, queue), ey(n);
        int m = 3;
        for (int i = 0; i < n; ++i) {
            ey[i] = m;
            m = std::max(m, std::abs(x[i]));
        }
        CHECK(hy == ey);
        // binary transform of a counting range and an array
        bc::transform(bc::make_counting_iterator(0), bc::make_counting_iterator(n), dx.begin(), dy.begin(),
                      bc::multiplies<int>(), queue);
        hy = host(dy, queue);
        bool ok = true;
        for (int i = 0; i < n; ++i) ok = ok && hy[i] == (int)((unsigned)i * (unsigned)x[i]);
        CHECK(ok);
        // accumulate over a mapped range
        const int s = bc::accumulate(bc::make_transform_iterator(dx.begin(), _1 - 10),
                                     bc::make_transform_iterator(dx.end(), _1 - 10), 1, queue);
        long es = 1;
        for (int i = 0; i < n; ++i) es += x[i] - 10;
        CHECK(s == (int)es);
    }
    std::printf("%d failures\n", failures);
    return failures;
}
