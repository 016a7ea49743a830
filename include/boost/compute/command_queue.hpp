//---------------------------------------------------------------------------//
// command_queue: an in-order queue = a bcmi_queue (one HIP stream plus its
// scratch arena and look-back state).  Copies share the queue (reference
// semantics: command_queue.hpp copies retain the same cl_command_queue).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_COMMAND_QUEUE_HPP
#define BOOST_COMPUTE_COMMAND_QUEUE_HPP

#include <memory>

#include <boost/compute/context.hpp>
#include <boost/compute/detail/bcmi.hpp>

namespace boost {
namespace compute {

class command_queue
{
public:
    command_queue() {}

    command_queue(const context &ctx, const device &dev) : m_context(ctx), m_device(dev)
    {
        bcmi_queue q = 0;
        detail::bcmi_check(bcmi_queue_create(dev.id(), &q));
        m_queue.reset(q, [](bcmi_queue p) { bcmi_queue_destroy(p); });
    }

    explicit command_queue(const device &dev) : command_queue(context(dev), dev) {}

    bcmi_queue get() const { return m_queue.get(); }
    const device &get_device() const { return m_device; }
    const context &get_context() const { return m_context; }

    // command_queue::finish (command_queue.hpp:1564)
    void finish() { detail::bcmi_check(bcmi_queue_finish(get())); }
    void flush() {}

    bool operator==(const command_queue &o) const { return m_queue == o.m_queue; }
    bool operator!=(const command_queue &o) const { return !(*this == o); }

private:
    context m_context;
    device m_device;
    std::shared_ptr<bcmi_queue_s> m_queue;
};

}  // namespace compute
}  // namespace boost

#endif
