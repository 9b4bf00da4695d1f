// Definitions of the opaque ABI handles and helpers shared by the library's translation units.
#pragma once
#include <string>
#include <vector>

#include "../../include/jagua_amd.h"
#include "jg_device.hpp"
#include "jg_host.hpp"

namespace jg {

struct DevBlob {                       // one device allocation, grown on demand
    void* ptr = nullptr;
    size_t cap = 0;
    DevBlob() = default;
    DevBlob(const DevBlob&) = delete;
    DevBlob& operator=(const DevBlob&) = delete;
    ~DevBlob();
    jg_status reserve(size_t bytes);
};

jg_status fail(jg_status code, const std::string& msg);
jg_status hip_check(hipError_t e, const char* what);
jg_status device_ready();

}  // namespace jg

struct jg_spolygon {
    jg::Polygon poly;
    jg::DevBlob dev;                   // vertices, ff poles, piers
    jg::ItemView view{};
    bool dev_valid = false;
    uint64_t upload_gen = 0;           // bumped on every upload (a batched plan's table refers to the upload)
};

struct jg_cde {
    jg_cde_config cfg{};
    jg::SlotMap map;
    jg::QuadTree qt;
    jg_hazkey exterior{0, 0};
    bool dirty = true;
    bool dev_valid = false;
    jg::DevBlob dev;                   // the device image (pool order, with headroom)
    jg::PoolImage img;                 // its host mirror
    jg::DevBlob staging_dev;           // patch segments and their bytes
    std::vector<uint8_t> staging;
    std::vector<jg::ImageSeg> segs;
    jg::DevBlob filter_dev;            // JG_FILTER_KEYS mask
    hipEvent_t filter_read = nullptr;  // completion of the last launch that read filter_dev
    jg::CdeView view{};
    jg_cde_stats stats{};
    uint64_t generation = 0;           // bumped on every upload of a new image
    ~jg_cde() {
        if (filter_read) (void)hipEventDestroy(filter_read);
    }
};

struct jg_snapshot {
    std::vector<jg::HazardRec> dynamic;
};

namespace jg {
jg_status cde_upload(jg_cde* c);
jg_status item_upload(jg_spolygon* it);
// the device record of a query item: its size, and its bytes written for device address dev
size_t item_image_bytes(const Polygon& p);
void item_image_write(const Polygon& p, uint8_t* host, uint8_t* dev, ItemView& view);
jg_status filter_view(jg_cde* c, const jg_hazard_filter* f, FilterView& out, hipStream_t stream);
jg_status filter_read_done(jg_cde* c, hipStream_t stream);
jg_status plan_prepare(jg_plan* p);
jg_status plan_enqueue(jg_plan* p, const float* d_m6, uint8_t* d_status, hipStream_t stream);
// Piers of many polygons on the GPU (generate_piers, fail_fast/piers.rs:25-98), bit for bit
// with the host routine; poles[i] are polygon i's relevant poles.  host[i] is set where the
// device could not finish the search the way the reference would; those take the host routine.
jg_status piers_batch(const std::vector<const Polygon*>& polys, const std::vector<std::vector<float4>>& poles,
                      uint32_t n_piers, std::vector<std::vector<float4>>& out, std::vector<uint8_t>& host);
}  // namespace jg
