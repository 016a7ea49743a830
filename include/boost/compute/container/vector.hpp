//---------------------------------------------------------------------------//
// vector<T>: a device array (reference: container/vector.hpp:103-823).  The
// constructors, iterators and element access the hot-path algorithms and
// their callers use; storage is one buffer (hipMalloc) owned by the vector.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_CONTAINER_VECTOR_HPP
#define BOOST_COMPUTE_CONTAINER_VECTOR_HPP

#include <algorithm>
#include <iterator>
#include <type_traits>
#include <vector>

#include <boost/compute/algorithm/copy.hpp>
#include <boost/compute/algorithm/fill.hpp>
#include <boost/compute/system.hpp>

namespace boost {
namespace compute {

template <class T>
class vector
{
public:
    typedef T value_type;
    typedef size_t size_type;
    typedef std::ptrdiff_t difference_type;
    typedef buffer_iterator<T> iterator;
    typedef buffer_iterator<T> const_iterator;

    // blocking element proxy (vector[i], reference: detail::buffer_value)
    class reference
    {
    public:
        reference(const buffer &b, size_t i) : m_it(b, i) {}
        operator T() const { return m_it.read(system::default_queue()); }
        reference &operator=(const T &v)
        {
            m_it.write(v, system::default_queue());
            return *this;
        }
        reference &operator=(const reference &o) { return *this = (T)o; }

    private:
        buffer_iterator<T> m_it;
    };

    explicit vector(const context &ctx = system::default_context()) : m_context(ctx), m_size(0) {}

    explicit vector(size_type count, const context &ctx = system::default_context())
        : m_context(ctx), m_size(count), m_buffer(ctx, count * sizeof(T))
    {
    }

    vector(size_type count, const T &value, command_queue &queue = system::default_queue())
        : m_context(queue.get_context()), m_size(count), m_buffer(queue.get_context(), count * sizeof(T))
    {
        ::boost::compute::fill(begin(), end(), value, queue);
    }

    template <class InputIterator,
              class = typename std::enable_if<!std::is_integral<InputIterator>::value>::type>
    vector(InputIterator first, InputIterator last, command_queue &queue = system::default_queue())
        : m_context(queue.get_context()), m_size((size_type)std::distance(first, last)),
          m_buffer(queue.get_context(), m_size * sizeof(T))
    {
        ::boost::compute::copy(first, last, begin(), queue);
    }

    vector(const std::vector<T> &v, command_queue &queue = system::default_queue())
        : vector(v.begin(), v.end(), queue)
    {
    }

    vector(const vector &other)
        : m_context(other.m_context), m_size(other.m_size), m_buffer(other.m_context, other.m_size * sizeof(T))
    {
        command_queue &q = system::default_queue();
        ::boost::compute::copy(other.begin(), other.end(), begin(), q);
        q.finish();
    }

    vector(vector &&other) noexcept
        : m_context(other.m_context), m_size(other.m_size), m_buffer(std::move(other.m_buffer))
    {
        other.m_size = 0;
    }

    vector &operator=(const vector &other)
    {
        if (this != &other) {
            vector tmp(other);
            swap(tmp);
        }
        return *this;
    }

    vector &operator=(vector &&other) noexcept
    {
        swap(other);
        return *this;
    }

    void swap(vector &o)
    {
        std::swap(m_context, o.m_context);
        std::swap(m_size, o.m_size);
        std::swap(m_buffer, o.m_buffer);
    }

    size_type size() const { return m_size; }
    bool empty() const { return m_size == 0; }
    size_type capacity() const { return m_buffer.size() / sizeof(T); }
    iterator begin() const { return iterator(m_buffer, 0); }
    iterator end() const { return iterator(m_buffer, m_size); }
    const buffer &get_buffer() const { return m_buffer; }
    const context &get_context() const { return m_context; }

    reference operator[](size_type i) { return reference(m_buffer, i); }
    T operator[](size_type i) const { return iterator(m_buffer, i).read(system::default_queue()); }

    void resize(size_type size, command_queue &queue = system::default_queue())
    {
        if (size <= capacity()) {
            m_size = size;
            return;
        }
        buffer nb(m_context, std::max<size_type>(size, 2 * capacity()) * sizeof(T));
        if (m_size) {
            detail::bcmi_check(bcmi_memcpy(queue.get(), nb.get(), m_buffer.get(), m_size * sizeof(T), BCMI_D2D, 0));
            queue.finish();
        }
        m_buffer = nb;
        m_size = size;
    }

    void push_back(const T &value, command_queue &queue = system::default_queue())
    {
        resize(m_size + 1, queue);
        iterator(m_buffer, m_size - 1).write(value, queue);
    }

    void clear() { m_size = 0; }

private:
    context m_context;
    size_type m_size;
    buffer m_buffer;
};

}  // namespace compute
}  // namespace boost

#endif
