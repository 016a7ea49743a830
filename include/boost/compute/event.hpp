//---------------------------------------------------------------------------//
// event (reference: event.hpp): a marker recorded on a queue; wait() blocks
// until the work enqueued before it has finished.  Copies share the
// underlying HIP event.
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_EVENT_HPP
#define BOOST_COMPUTE_EVENT_HPP

#include <memory>

#include <bcmi.h>

#include <boost/compute/command_queue.hpp>
#include <boost/compute/detail/bcmi.hpp>

namespace boost {
namespace compute {

class event {
public:
    event() {}
    // records an event at the current end of the queue
    static event marker(command_queue &queue)
    {
        bcmi_event e = 0;
        detail::bcmi_check(bcmi_event_create(&e));
        event ev;
        ev.m_event.reset(e, [](bcmi_event p) { bcmi_event_destroy(p); });
        detail::bcmi_check(bcmi_event_record(e, queue.get()));
        return ev;
    }
    void wait() const
    {
        if (m_event) detail::bcmi_check(bcmi_event_synchronize(m_event.get()));
    }
    bool valid() const { return (bool)m_event; }
    bcmi_event get() const { return m_event.get(); }

private:
    std::shared_ptr<bcmi_event_s> m_event;
};

}  // namespace compute
}  // namespace boost

#endif
