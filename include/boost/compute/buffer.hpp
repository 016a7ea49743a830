//---------------------------------------------------------------------------//
// buffer: a reference-counted device allocation (reference: buffer.hpp:91-121,
// copies share the handle).  Allocation goes through bcmi_malloc (hipMalloc).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_BUFFER_HPP
#define BOOST_COMPUTE_BUFFER_HPP

#include <memory>

#include <boost/compute/context.hpp>
#include <boost/compute/detail/bcmi.hpp>

namespace boost {
namespace compute {

class buffer
{
public:
    buffer() : m_size(0) {}

    buffer(const context &ctx, size_t size) : m_context(ctx), m_size(size)
    {
        void *p = 0;
        if (size) detail::bcmi_check(bcmi_malloc(size, &p));
        m_mem.reset(p, [](void *q) { bcmi_free(q); });
    }

    void *get() const { return m_mem.get(); }
    size_t size() const { return m_size; }
    const context &get_context() const { return m_context; }

    bool operator==(const buffer &o) const { return m_mem == o.m_mem; }
    bool operator!=(const buffer &o) const { return !(*this == o); }

private:
    context m_context;
    std::shared_ptr<void> m_mem;
    size_t m_size;
};

}  // namespace compute
}  // namespace boost

#endif
