// SisterMultiviewDisparities.hpp -- the reference's helper class
// (cpp/include/sister/SisterMultiviewDisparities.hpp:18-26) on top of the C ABI
// in include/sister.h, computing on one MI355X.
//
//   SisterMultiviewDisparities(center, right, top, left, bottom)
//   void compute_disparities(int dispCount, Mat16& multiview, Mat16& horizontal,
//                            Mat16& vertical)
//
// Same argument order and meaning: five views in the constructor order of the
// reference (hpp:22), 8-bit gray or 3-channel BGR (converted like
// cvtColor(BGR2GRAY), hpp:29-33), outputs W x H uint16 maps holding
// min(255*d, 65535) (hpp:111-118).  Where OpenCV is available the cv::Mat
// signatures of the reference are provided as well.
//
// Errors: the reference lets OpenCV throw cv::Exception; this class throws
// sister::Error (a std::runtime_error) carrying the library's status and
// message.  The C ABI underneath never throws.
#ifndef SISTER_MULTIVIEW_DISPARITIES_HPP
#define SISTER_MULTIVIEW_DISPARITIES_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sister.h"

#if defined(__has_include)
#if __has_include(<opencv2/core.hpp>) && __has_include(<opencv2/imgproc.hpp>)
#include <opencv2/core.hpp>
#include <opencv2/imgproc.hpp>
#define SISTER_HAVE_OPENCV 1
#endif
#endif

namespace sister {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Minimal owning image: rows x cols x channels, row-major, tightly packed.
template <typename T>
struct Image {
    int rows = 0, cols = 0, channels = 1;
    std::vector<T> data;
    Image() = default;
    Image(int r, int c, int ch = 1) : rows(r), cols(c), channels(ch), data((size_t)r * c * ch) {}
    bool empty() const { return data.empty(); }
    T* ptr(int r) { return data.data() + (size_t)r * cols * channels; }
    const T* ptr(int r) const { return data.data() + (size_t)r * cols * channels; }
};
using Mat8 = Image<uint8_t>;
using Mat16 = Image<uint16_t>;

// cvtColor(COLOR_BGR2GRAY) for 8-bit input: fixed-point weights 1868 (B),
// 9617 (G), 4899 (R), 14-bit rounding; gray input is returned unchanged.
inline Mat8 to_gray(const Mat8& in)
{
    if (in.channels == 1) return in;
    if (in.channels < 3) throw Error(SISTER_EINVAL, "expected 1, 3 or 4 channels");
    Mat8 g(in.rows, in.cols, 1);
    for (int r = 0; r < in.rows; r++) {
        const uint8_t* s = in.ptr(r);
        uint8_t* d = g.ptr(r);
        for (int c = 0; c < in.cols; c++, s += in.channels)
            d[c] = (uint8_t)((s[0] * 1868u + s[1] * 9617u + s[2] * 4899u + (1u << 13)) >> 14);
    }
    return g;
}

// one context per process by default; pass your own for several devices
class Context {
public:
    explicit Context(int device = 0, size_t hbm_budget = 0)
    {
        sister_config cfg{device, 0, hbm_budget};
        int rc = sister_create(&cfg, &ctx_);
        if (rc != SISTER_OK) throw Error(rc, "sister_create failed (no HIP device?)");
    }
    ~Context() { sister_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    sister_ctx* get() const { return ctx_; }
    void check(int rc) const
    {
        if (rc != SISTER_OK) throw Error(rc, sister_last_error(ctx_));
    }
    static Context& shared()
    {
        static Context c;
        return c;
    }

private:
    sister_ctx* ctx_ = nullptr;
};

}  // namespace sister

class SisterMultiviewDisparities {
public:
    SisterMultiviewDisparities(const sister::Mat8& center, const sister::Mat8& right, const sister::Mat8& top,
                               const sister::Mat8& left, const sister::Mat8& bottom,
                               sister::Context* ctx = nullptr)
        : center(center), right(right), top(top), left(left), bottom(bottom), ctx_(ctx)
    {
    }

    // hpp:26-119: all three maps (modes 0, 1, 2)
    void compute_disparities(int dispCount, sister::Mat16& disp_multiview, sister::Mat16& disp_horizontal,
                             sister::Mat16& disp_vertical)
    {
        compute(dispCount, SISTER_MODE_ALL, &disp_multiview, &disp_horizontal, &disp_vertical, nullptr);
    }

    // the multiview map alone, optionally with the sub-pixel refinement
    void compute_multiview(int dispCount, sister::Mat16& disp_multiview, sister::Image<float>* subpix = nullptr)
    {
        compute(dispCount, SISTER_MODE_MULTIVIEW, &disp_multiview, nullptr, nullptr, subpix);
    }

#ifdef SISTER_HAVE_OPENCV
    SisterMultiviewDisparities(cv::Mat c, cv::Mat r, cv::Mat t, cv::Mat l, cv::Mat b, sister::Context* ctx = nullptr)
        : center(from_cv(c)), right(from_cv(r)), top(from_cv(t)), left(from_cv(l)), bottom(from_cv(b)), ctx_(ctx)
    {
    }
    void compute_disparities(int dispCount, cv::Mat& disp_multiview, cv::Mat& disp_horizontal, cv::Mat& disp_vertical)
    {
        sister::Mat16 m, h, v;
        compute_disparities(dispCount, m, h, v);
        disp_multiview = to_cv(m);
        disp_horizontal = to_cv(h);
        disp_vertical = to_cv(v);
    }
#endif

private:
    void compute(int dispCount, int modes, sister::Mat16* m, sister::Mat16* h, sister::Mat16* v,
                 sister::Image<float>* sp)
    {
        sister::Context& ctx = ctx_ ? *ctx_ : sister::Context::shared();
        const sister::Mat8* in[5] = {&center, &right, &top, &left, &bottom};
        sister::Mat8 gray[5];
        for (int i = 0; i < 5; i++) {
            if (in[i]->empty()) throw sister::Error(SISTER_EINVAL, "empty input view");
            gray[i] = sister::to_gray(*in[i]);
            if (gray[i].rows != gray[0].rows || gray[i].cols != gray[0].cols)
                throw sister::Error(SISTER_EINVAL, "the five views must have the same size");
        }
        const int W = gray[0].cols, H = gray[0].rows;
        sister_rig r{};
        r.width = W;
        r.height = H;
        r.stride = W;
        r.disp_count = dispCount;
        for (int i = 0; i < 5; i++) r.view[i] = gray[i].data.data();
        if (m) { *m = sister::Mat16(H, W); r.out_multiview = m->data.data(); }
        if (h) { *h = sister::Mat16(H, W); r.out_horizontal = h->data.data(); }
        if (v) { *v = sister::Mat16(H, W); r.out_vertical = v->data.data(); }
        if (sp) { *sp = sister::Image<float>(H, W); r.out_subpix = sp->data.data(); }
        ctx.check(sister_compute_rig(ctx.get(), &r, modes));
    }

#ifdef SISTER_HAVE_OPENCV
    static sister::Mat8 from_cv(const cv::Mat& a)
    {
        cv::Mat g;
        if (a.channels() == 1) g = a; else cv::cvtColor(a, g, cv::COLOR_BGR2GRAY);
        sister::Mat8 out(g.rows, g.cols, 1);
        for (int r = 0; r < g.rows; r++) std::copy(g.ptr<uint8_t>(r), g.ptr<uint8_t>(r) + g.cols, out.ptr(r));
        return out;
    }
    static cv::Mat to_cv(const sister::Mat16& a)
    {
        cv::Mat out(a.rows, a.cols, CV_16UC1);
        for (int r = 0; r < a.rows; r++) std::copy(a.ptr(r), a.ptr(r) + a.cols, out.ptr<uint16_t>(r));
        return out;
    }
#endif

    sister::Mat8 center, right, top, left, bottom;
    sister::Context* ctx_;
};

#endif  // SISTER_MULTIVIEW_DISPARITIES_HPP
