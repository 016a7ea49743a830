//---------------------------------------------------------------------------//
// device: a HIP device seen through the C-ABI (reference: device.hpp).
// Only gfx950 GPUs are served, so type() is always gpu: the algorithms take
// the reference's GPU column (e.g. sort.hpp:107-122).
//---------------------------------------------------------------------------//
#ifndef BOOST_COMPUTE_DEVICE_HPP
#define BOOST_COMPUTE_DEVICE_HPP

#include <string>

#include <boost/compute/detail/bcmi.hpp>

namespace boost {
namespace compute {

class device
{
public:
    enum type_t { cpu = 1 << 1, gpu = 1 << 2, accelerator = 1 << 3 };

    device() : m_id(-1) {}
    explicit device(int id) : m_id(id) {}

    int id() const { return m_id; }
    int get() const { return m_id; }
    unsigned type() const { return gpu; }

    std::string name() const { return info().name; }
    std::string arch() const { return info().arch; }
    unsigned compute_units() const { return (unsigned)info().compute_units; }
    size_t global_memory_size() const { return info().global_mem_bytes; }
    size_t local_memory_size() const { return info().lds_bytes_per_block; }

    bool operator==(const device &o) const { return m_id == o.m_id; }
    bool operator!=(const device &o) const { return m_id != o.m_id; }

private:
    bcmi_device_info info() const
    {
        bcmi_device_info i;
        detail::bcmi_check(bcmi_device_get_info(m_id, &i));
        return i;
    }

    int m_id;
};

}  // namespace compute
}  // namespace boost

#endif
