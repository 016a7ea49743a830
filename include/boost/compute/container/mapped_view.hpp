//---------------------------------------------------------------------------//
// mapped_view<T> (reference: container/mapped_view.hpp:43-244): a device
// range over a host array, for algorithms on host data (the reference's
// host-iterator sort maps the host range this way, sort.hpp:124-148).
//
// The reference wraps the host memory in a CL_MEM_USE_HOST_PTR buffer, so the
// device works on it in place and map() / unmap() synchronise the two sides.
// Here the view owns a device copy of the host array: construction uploads it,
// map() downloads it (unless the mapping only invalidates the region) so the
// host sees the device's results, and unmap() uploads the host array again
// when the mapping allowed writes.  map() and unmap() block, as the
// reference's blocking map does.  Algorithms take begin()/end() like any
// device range, on any queue.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_CONTAINER_MAPPED_VIEW_HPP
#define BOOST_COMPUTE_CONTAINER_MAPPED_VIEW_HPP

#include <cstddef>
#include <cstdint>

#include <bcmi.h>

#include <boost/compute/buffer.hpp>
#include <boost/compute/command_queue.hpp>
#include <boost/compute/context.hpp>
#include <boost/compute/detail/bcmi.hpp>
#include <boost/compute/iterator/buffer_iterator.hpp>
#include <boost/compute/system.hpp>

#ifndef CL_MAP_READ
typedef uint64_t cl_map_flags;
#define CL_MAP_READ (1 << 0)
#define CL_MAP_WRITE (1 << 1)
#define CL_MAP_WRITE_INVALIDATE_REGION (1 << 2)
#endif

namespace boost {
namespace compute {

template <class T>
class mapped_view
{
public:
    typedef T value_type;
    typedef std::size_t size_type;
    typedef std::ptrdiff_t difference_type;
    typedef buffer_iterator<T> iterator;
    typedef buffer_iterator<T> const_iterator;

    mapped_view() : m_host(0), m_size(0), m_read_only(false), m_flags(0) {}

    mapped_view(T *host_ptr, size_type size, const context &ctx = system::default_context())
        : m_host(host_ptr), m_size(size), m_read_only(false), m_flags(0), m_buffer(ctx, size * sizeof(T))
    {
        upload(system::default_queue());
    }

    // a read-only view: unmap() never writes back to the host array
    mapped_view(const T *host_ptr, size_type size, const context &ctx = system::default_context())
        : m_host(const_cast<T *>(host_ptr)), m_size(size), m_read_only(true), m_flags(0),
          m_buffer(ctx, size * sizeof(T))
    {
        upload(system::default_queue());
    }

    iterator begin() { return iterator(m_buffer, 0); }
    const_iterator begin() const { return const_iterator(m_buffer, 0); }
    const_iterator cbegin() const { return begin(); }
    iterator end() { return iterator(m_buffer, m_size); }
    const_iterator end() const { return const_iterator(m_buffer, m_size); }
    const_iterator cend() const { return end(); }

    size_type size() const { return m_size; }
    T *get_host_ptr() { return m_host; }
    const T *get_host_ptr() const { return m_host; }
    const buffer &get_buffer() const { return m_buffer; }

    // a view of the first `size` elements of the same host array (which must
    // hold that many), uploaded afresh
    void resize(size_type size)
    {
        m_buffer = buffer(m_buffer.get_context(), size * sizeof(T));
        m_size = size;
        upload(system::default_queue());
    }

    void map(cl_map_flags flags, command_queue &queue)
    {
        m_flags = flags;
        if (flags != CL_MAP_WRITE_INVALIDATE_REGION && m_size)
            detail::bcmi_check(bcmi_memcpy(queue.get(), m_host, m_buffer.get(), m_size * sizeof(T), BCMI_D2H, 1));
    }

    void map(command_queue &queue) { map(CL_MAP_READ | CL_MAP_WRITE, queue); }

    void unmap(command_queue &queue)
    {
        if (!m_read_only && (m_flags & (CL_MAP_WRITE | CL_MAP_WRITE_INVALIDATE_REGION))) upload(queue);
        m_flags = 0;
    }

private:
    void upload(command_queue &queue)
    {
        if (m_size)
            detail::bcmi_check(bcmi_memcpy(queue.get(), m_buffer.get(), m_host, m_size * sizeof(T), BCMI_H2D, 1));
    }

    T *m_host;
    size_type m_size;
    bool m_read_only;
    cl_map_flags m_flags;
    buffer m_buffer;
};

}  // namespace compute
}  // namespace boost

#endif
