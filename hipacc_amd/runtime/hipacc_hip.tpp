// hipacc_hip.tpp -- templates of the MI355X host runtime: memory, copies,
// operator launches, pyramids.  Included by hipacc_hip.hpp.
#ifndef HIPACC_HIP_TPP
#define HIPACC_HIP_TPP

// ---------------------------------------------------------------------------
// Memory (runtime/hipacc_cu.tpp:33-193)
// ---------------------------------------------------------------------------
template <typename T>
void hipaccWriteMemory(HipaccImageHip<T>& img, const T* host_mem);

// Wraps a device allocation; uploads host_mem (or the zeroed host mirror).
template <typename T>
HipaccImageHip<T> createImage(const T* host_mem, T* mem, int width, int height, int stride,
                              int alignment) {
  HipaccImageHip<T> img =
      std::make_shared<HipaccImageHipRaw<T>>(width, height, stride, alignment, mem);
  hipaccWriteMemory<T>(img, host_mem ? host_mem : img->get_host_memory());
  return img;
}

// Pitched image: row pitch a multiple of `alignment` bytes (rounded up to a
// multiple of the pixel size), as hipaccCreateMemory with alignment.
template <typename T>
HipaccImageHip<T> hipaccCreateMemory(const T* host_mem, int width, int height, int alignment) {
  alignment = (alignment + (int)sizeof(T) - 1) / (int)sizeof(T) * (int)sizeof(T);
  void* mem = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&mem, &stride, width, height, (int)sizeof(T), alignment),
                  "hipaccCreateMemory"))
    return nullptr;
  return createImage<T>(host_mem, static_cast<T*>(mem), width, height, stride, alignment);
}

// Dense image (stride == width), as hipaccCreateMemory without alignment.
template <typename T>
HipaccImageHip<T> hipaccCreateMemory(const T* host_mem, int width, int height) {
  void* mem = nullptr;
  int32_t stride = 0;
  if (hipaccCheck(hipacc_hip_alloc2d(&mem, &stride, width, height, (int)sizeof(T),
                                     (int)sizeof(T)),
                  "hipaccCreateMemory"))
    return nullptr;
  return createImage<T>(host_mem, static_cast<T*>(mem), width, height, stride, 0);
}

// Host -> device; host_mem is dense (width pixels per row) and is also kept
// in the image's host mirror.
template <typename T>
void hipaccWriteMemory(HipaccImageHip<T>& img, const T* host_mem) {
  if (!img || host_mem == nullptr) return;
  const size_t n = (size_t)img->get_width() * img->get_height();
  if (host_mem != img->get_host_memory())
    std::copy(host_mem, host_mem + n, img->get_host_memory());
  hipaccCheck(hipacc_hip_upload2d(img->get_device_memory(), img->get_stride(), host_mem,
                                  img->get_width(), img->get_height(), (int)sizeof(T), nullptr),
              "hipaccWriteMemory");
}

// Device -> host mirror; returns the mirror (owned by the image).
template <typename T>
T* hipaccReadMemory(const HipaccImageHip<T>& img) {
  auto* raw = const_cast<HipaccImageHipBase<T>*>(img.get());
  hipaccCheck(hipacc_hip_download2d(raw->get_host_memory(), raw->get_device_memory(),
                                    raw->get_stride(), raw->get_width(), raw->get_height(),
                                    (int)sizeof(T), nullptr),
              "hipaccReadMemory");
  return raw->get_host_memory();
}

template <typename T>
HipaccImageHip<T> hipaccMapMemory(HipaccImageHip<T> img) { return img; }

// Device-to-device copy of a whole image / of accessor windows
// (runtime/hipacc_cu_standalone.hpp:166-216).
template <typename T>
void hipaccCopyMemoryRegion(const HipaccAccessor<T>& src, const HipaccAccessor<T>& dst,
                            const HipaccExecutionParameterHip& ep = nullptr) {
  hipacc_view s = hipaccView(src), d = hipaccView(dst);
  hipaccCheck(hipacc_hip_copy_region(&s, &d, hipaccStreamOf(ep)), "hipaccCopyMemoryRegion");
}
template <typename T>
void hipaccCopyMemory(const HipaccImageHip<T>& src, HipaccImageHip<T>& dst) {
  hipaccCopyMemoryRegion(HipaccAccessor<T>(src), HipaccAccessor<T>(dst));
}

// ---------------------------------------------------------------------------
// Operator launches.  Each returns the library status (0 = OK) after
// logging a failure.
// ---------------------------------------------------------------------------
template <typename M> struct hipacc_mask_acc { static constexpr int value = HIPACC_I32; };
template <> struct hipacc_mask_acc<float> { static constexpr int value = HIPACC_F32; };

// A convolve()/reduce()/iterate() kernel over `mask` (coefficients and
// domain): the generated local-operator kernel of lib/AST/Convolution.cpp.
// The accumulator type follows the mask: float mask -> float lambda,
// integer mask -> int lambda.
template <typename TI, typename TO, typename M>
int hipaccLaunchLocal(const HipaccAccessor<TI>& in, const HipaccAccessor<TO>& out,
                      const HipaccMask<M>& mask, const HipaccLocalOptions& opt = {},
                      const HipaccExecutionParameterHip& ep = nullptr) {
  std::vector<float> wf;
  std::vector<int32_t> wi;
  hipacc_local_desc d;
  std::memset(&d, 0, sizeof(d));
  d.in = hipaccView(in);
  d.out = hipaccView(out);
  d.mask_w = mask.size_x;
  d.mask_h = mask.size_y;
  d.acc_dtype = hipacc_mask_acc<M>::value;
  if (d.acc_dtype == HIPACC_F32) {
    wf.assign(mask.coef.begin(), mask.coef.end());
    d.weights_f32 = wf.data();
  } else {
    wi.assign(mask.coef.begin(), mask.coef.end());
    d.weights_i32 = wi.data();
  }
  d.domain = mask.domain.data();
  d.use_weights = opt.use_weights ? 1 : 0;
  d.reduce = (int32_t)opt.reduce;
  d.boundary = (int32_t)opt.boundary;
  d.const_val = opt.const_val;
  d.cvt = opt.round ? HIPACC_CVT_ROUND : HIPACC_CVT_TRUNC;
  d.post_div = opt.post_div;
  d.fp_contract = opt.fp_contract ? 1 : 0;
  if (ep) ep->pre_kernel();
  int st = hipaccCheck(hipacc_hip_local(&d, hipaccStreamOf(ep)), "hipaccLaunchLocal");
  if (ep) ep->post_kernel();
  return st;
}

// A row mask then a column mask (the two-kernel separable program, one
// launch).  Float masks accumulate in float, integer masks in int32.
template <typename TI, typename TO, typename M>
int hipaccLaunchSeparable(const HipaccAccessor<TI>& in, const HipaccAccessor<TO>& out,
                          const std::vector<M>& row, const std::vector<M>& col,
                          const HipaccLocalOptions& opt = {},
                          const HipaccExecutionParameterHip& ep = nullptr) {
  std::vector<float> rf, cf;
  std::vector<int32_t> ri, ci;
  hipacc_separable_desc d;
  std::memset(&d, 0, sizeof(d));
  d.in = hipaccView(in);
  d.out = hipaccView(out);
  d.ksize_x = (int32_t)row.size();
  d.ksize_y = (int32_t)col.size();
  d.acc_dtype = hipacc_mask_acc<M>::value;
  if (d.acc_dtype == HIPACC_F32) {
    rf.assign(row.begin(), row.end());
    cf.assign(col.begin(), col.end());
    d.row_f32 = rf.data();
    d.col_f32 = cf.data();
  } else {
    ri.assign(row.begin(), row.end());
    ci.assign(col.begin(), col.end());
    d.row_i32 = ri.data();
    d.col_i32 = ci.data();
  }
  d.boundary = (int32_t)opt.boundary;
  d.const_val = opt.const_val;
  d.cvt = opt.round ? HIPACC_CVT_ROUND : HIPACC_CVT_TRUNC;
  d.post_div = opt.post_div;
  d.fp_contract = opt.fp_contract ? 1 : 0;
  if (ep) ep->pre_kernel();
  int st = hipaccCheck(hipacc_hip_separable(&d, hipaccStreamOf(ep)), "hipaccLaunchSeparable");
  if (ep) ep->post_kernel();
  return st;
}

// A point operator (hipacc_point_op) over one or two images.
template <typename T0, typename TO, typename T1 = T0>
int hipaccLaunchPoint(int op, const HipaccAccessor<T0>& in0, const HipaccAccessor<TO>& out,
                      const HipaccAccessor<T1>* in1 = nullptr, float alpha = 1.0f,
                      float beta = 0.0f, bool round = false,
                      const HipaccExecutionParameterHip& ep = nullptr) {
  hipacc_point_desc d;
  std::memset(&d, 0, sizeof(d));
  d.in0 = hipaccView(in0);
  if (in1) d.in1 = hipaccView(*in1);
  d.out = hipaccView(out);
  d.op = op;
  d.alpha = alpha;
  d.beta = beta;
  d.cvt = round ? HIPACC_CVT_ROUND : HIPACC_CVT_TRUNC;
  if (ep) ep->pre_kernel();
  int st = hipaccCheck(hipacc_hip_point(&d, hipaccStreamOf(ep)), "hipaccLaunchPoint");
  if (ep) ep->post_kernel();
  return st;
}

// One stage of a fused chain (hipacc_hip_fused): a point operator on the
// previous image (and, for a two-input op, on in1) or a local operator over
// it, writing pixels of type dtype.
struct HipaccFusedStage {
  int kind = HIPACC_STAGE_POINT;
  int dtype = HIPACC_F32;
  int op = HIPACC_POINT_COPY;
  float alpha = 1.0f, beta = 0.0f;
  bool round = false;
  hipacc_view in1{};
  int mask_w = 0, mask_h = 0, acc_dtype = HIPACC_F32;
  std::vector<float> wf;
  std::vector<int32_t> wi;
  std::vector<uint8_t> domain;
  HipaccLocalOptions opt;
};

// A one-input point stage writing pixels of type TO.
template <typename TO>
HipaccFusedStage hipaccFusedPoint(int op, float alpha = 1.0f, float beta = 0.0f,
                                  bool round = false) {
  HipaccFusedStage s;
  s.dtype = hipacc_pixel_code<TO>::value;
  s.op = op;
  s.alpha = alpha;
  s.beta = beta;
  s.round = round;
  return s;
}

// A two-input point stage (ADD, SUB, ...) with in1 as its second operand.
template <typename TO, typename T1>
HipaccFusedStage hipaccFusedPoint2(int op, const HipaccAccessor<T1>& in1) {
  HipaccFusedStage s = hipaccFusedPoint<TO>(op);
  s.in1 = hipaccView(in1);
  return s;
}

// A local stage writing pixels of type TO; the accumulator follows the mask.
template <typename TO, typename M>
HipaccFusedStage hipaccFusedLocal(const HipaccMask<M>& mask, const HipaccLocalOptions& opt) {
  HipaccFusedStage s;
  s.kind = HIPACC_STAGE_LOCAL;
  s.dtype = hipacc_pixel_code<TO>::value;
  s.mask_w = mask.size_x;
  s.mask_h = mask.size_y;
  s.acc_dtype = hipacc_mask_acc<M>::value;
  if (s.acc_dtype == HIPACC_F32)
    s.wf.assign(mask.coef.begin(), mask.coef.end());
  else
    s.wi.assign(mask.coef.begin(), mask.coef.end());
  s.domain = mask.domain;
  s.opt = opt;
  return s;
}

// The chain of stages from `in` to `out` as one launch (the fused kernel of
// lib/AST/ASTFuse.cpp:212-413): bit-identical to launching the stages one
// by one through intermediate images of out's size.
template <typename TI, typename TO>
int hipaccLaunchFused(const HipaccAccessor<TI>& in, const HipaccAccessor<TO>& out,
                      const std::vector<HipaccFusedStage>& stages,
                      const HipaccExecutionParameterHip& ep = nullptr) {
  hipacc_fused_desc d;
  std::memset(&d, 0, sizeof(d));
  hipacc_local_desc locals[HIPACC_MAX_FUSED_STAGES];
  std::memset(locals, 0, sizeof(locals));
  if (stages.empty() || stages.size() > HIPACC_MAX_FUSED_STAGES)
    return hipaccCheck(HIPACC_ERR_INVALID_ARG, "hipaccLaunchFused: 1..8 stages");
  d.in = hipaccView(in);
  d.out = hipaccView(out);
  d.num_stages = (int32_t)stages.size();
  for (size_t k = 0; k < stages.size(); ++k) {
    const HipaccFusedStage& st = stages[k];
    hipacc_fused_stage& g = d.stages[k];
    g.kind = st.kind;
    g.dtype = st.dtype;
    if (st.kind == HIPACC_STAGE_POINT) {
      g.op = st.op;
      g.alpha = st.alpha;
      g.beta = st.beta;
      g.cvt = st.round ? HIPACC_CVT_ROUND : HIPACC_CVT_TRUNC;
      g.in1 = st.in1;
    } else {
      hipacc_local_desc& l = locals[k];
      l.mask_w = st.mask_w;
      l.mask_h = st.mask_h;
      l.acc_dtype = st.acc_dtype;
      l.weights_f32 = st.wf.empty() ? nullptr : st.wf.data();
      l.weights_i32 = st.wi.empty() ? nullptr : st.wi.data();
      l.domain = st.domain.data();
      l.use_weights = st.opt.use_weights ? 1 : 0;
      l.reduce = (int32_t)st.opt.reduce;
      l.boundary = (int32_t)st.opt.boundary;
      l.const_val = st.opt.const_val;
      l.cvt = st.opt.round ? HIPACC_CVT_ROUND : HIPACC_CVT_TRUNC;
      l.post_div = st.opt.post_div;
      l.fp_contract = st.opt.fp_contract ? 1 : 0;
      g.local = &l;
    }
  }
  if (ep) ep->pre_kernel();
  int st = hipaccCheck(hipacc_hip_fused(&d, hipaccStreamOf(ep)), "hipaccLaunchFused");
  if (ep) ep->post_kernel();
  return st;
}

// The bilateral filter kernel (range weight from sigma_r, spatial weight
// from sigma_d or a host table).
// The spatial table of a bilateral kernel written in the DSL (a Mask of
// weights iterated over a Domain): the mask's coefficients, 0 where the
// domain has a hole (a tap of weight 0 adds exactly 0 to both sums, as a
// skipped tap does).  Empty if the two are not one square window.
template <typename M, typename D>
std::vector<float> hipaccBilateralSpatial(const HipaccMask<M>& mask, const HipaccMask<D>& dom) {
  std::vector<float> w;
  if (mask.size_x != mask.size_y || dom.size_x != mask.size_x || dom.size_y != mask.size_y) {
    hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                            "ERROR: hipaccBilateralSpatial: mask and domain are not one square "
                            "window");
    return w;
  }
  w.resize(mask.coef.size());
  for (size_t i = 0; i < w.size(); ++i) w[i] = dom.domain[i] ? (float)mask.coef[i] : 0.0f;
  return w;
}

template <typename T>
int hipaccLaunchBilateral(const HipaccAccessor<T>& in, const HipaccAccessor<T>& out, int ksize,
                          float sigma_d, float sigma_r, const float* spatial = nullptr,
                          HipaccBoundary boundary = HipaccBoundary::CLAMP, double const_val = 0.0,
                          bool round = false, const HipaccExecutionParameterHip& ep = nullptr) {
  hipacc_bilateral_desc d;
  std::memset(&d, 0, sizeof(d));
  d.in = hipaccView(in);
  d.out = hipaccView(out);
  d.ksize = ksize;
  d.sigma_d = sigma_d;
  d.sigma_r = sigma_r;
  d.spatial = spatial;
  d.boundary = (int32_t)boundary;
  d.const_val = const_val;
  d.cvt = round ? HIPACC_CVT_ROUND : HIPACC_CVT_TRUNC;
  if (ep) ep->pre_kernel();
  int st = hipaccCheck(hipacc_hip_bilateral(&d, hipaccStreamOf(ep)), "hipaccLaunchBilateral");
  if (ep) ep->post_kernel();
  return st;
}

// ---------------------------------------------------------------------------
// Pyramids (runtime/hipacc_base.hpp:131-191, runtime/hipacc_cu.hpp:207-224,
// runtime/hipacc_cu.tpp:466-498)
// ---------------------------------------------------------------------------
class HipaccPyramid {
 protected:
  const int depth_;
  int level_ = 0;
  bool bound_ = false;

 public:
  explicit HipaccPyramid(int depth) : depth_(depth) {}
  virtual ~HipaccPyramid() = default;
  int depth() const { return depth_; }
  int level() const { return level_; }
  void levelInc() { ++level_; }
  void levelDec() { --level_; }
  bool is_top_level() const { return level_ == 0; }
  bool is_bottom_level() const { return level_ == depth_ - 1; }
  // Takes the pyramid for a traversal (level 0); false if already taken.
  bool bind() {
    if (bound_) return false;
    bound_ = true;
    level_ = 0;
    return true;
  }
  void unbind() { bound_ = false; }
};

template <typename T>
class HipaccPyramidHip final : public HipaccPyramid {
  std::vector<HipaccImageHip<T>> imgs_;

 public:
  explicit HipaccPyramidHip(int depth) : HipaccPyramid(depth) {}
  void add(const HipaccImageHip<T>& img) { imgs_.push_back(img); }
  // Image `relative` levels below the current one (-1: the finer level).
  HipaccImageHip<T> operator()(int relative) {
    const int l = level_ + relative;
    if (l < 0 || l >= (int)imgs_.size()) {
      hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                              "ERROR: pyramid level " + std::to_string(l) + " out of bounds");
      return nullptr;
    }
    return imgs_[l];
  }
  void swap(HipaccPyramidHip& other) { imgs_.swap(other.imgs_); }
};

template <typename T>
HipaccImageHip<T> hipaccCreatePyramidImage(const HipaccImageHip<T>& base, int width, int height) {
  if (base->get_alignment() > 0)
    return hipaccCreateMemory<T>(nullptr, width, height, base->get_alignment());
  return hipaccCreateMemory<T>(nullptr, width, height);
}

// Level 0 is `img`; level i is (w >> i) x (h >> i).
template <typename T>
HipaccPyramidHip<T> hipaccCreatePyramid(const HipaccImageHip<T>& img, int depth) {
  HipaccPyramidHip<T> p(depth);
  p.add(img);
  int w = img->get_width() / 2, h = img->get_height() / 2;
  for (int i = 1; i < depth; ++i) {
    if (w * h <= 0) {
      hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                              "ERROR: pyramid stages too deep for image size");
      break;
    }
    p.add(hipaccCreatePyramidImage<T>(img, w, h));
    w /= 2;
    h /= 2;
  }
  return p;
}

// Recursive traversal: hipaccTraverse(pyramids..., body) runs body at
// level 0; inside it, hipaccTraverse(loop, between) descends one level,
// runs the body `loop` times there (calling `between` between runs) and
// comes back up (runtime/hipacc_base_standalone.hpp:63-255).
class HipaccPyramidTraversor {
  std::vector<const std::function<void()>*> funcs_;
  std::vector<std::vector<HipaccPyramid*>> pyrs_;

 public:
  void pushFunc(const std::function<void()>* f) { funcs_.push_back(f); }
  void popFunc() { funcs_.pop_back(); }
  const std::function<void()> getLastFunc() { return *funcs_.back(); }
  void pushPyramids(std::vector<HipaccPyramid*>& p) { pyrs_.push_back(p); }
  void popPyramids() { pyrs_.pop_back(); }
  std::vector<HipaccPyramid*> getLastPyramids() { return pyrs_.back(); }
  bool hasPyramids() { return !pyrs_.empty(); }

  void hipaccTraverse(std::vector<HipaccPyramid*> pyrs, const std::function<void()>& func) {
    for (size_t i = 0; i < pyrs.size(); ++i) {
      if (i + 1 < pyrs.size() && pyrs[i]->depth() != pyrs[i + 1]->depth()) {
        hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                                "ERROR: pyramid depths do not match");
        return;
      }
    }
    size_t bound = 0;
    for (; bound < pyrs.size(); ++bound)
      if (!pyrs[bound]->bind()) break;
    if (bound != pyrs.size()) {
      for (size_t i = 0; i < bound; ++i) pyrs[i]->unbind();
      hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                              "ERROR: pyramid already bound to another traversal");
      return;
    }
    pushPyramids(pyrs);
    pushFunc(&func);
    getLastFunc()();
    popFunc();
    popPyramids();
    for (auto* p : pyrs) p->unbind();
  }
  // hipaccTraverse(p0[, p1, ...], body): 1 to any number of pyramids of
  // equal depth, then the body.
  template <typename... P>
  void hipaccTraverse(HipaccPyramid& p0, P&&... rest_and_func) {
    collect({&p0}, std::forward<P>(rest_and_func)...);
  }
  void hipaccTraverse(unsigned int loop = 1, const std::function<void()>& func = [] {}) {
    if (!hasPyramids()) {
      hipaccRuntimeLogTrivial(hipaccRuntimeLogLevel::ERROR,
                              "ERROR: traverse recursion called outside of traverse");
      return;
    }
    std::vector<HipaccPyramid*> pyrs = getLastPyramids();
    if (pyrs.at(0)->is_bottom_level()) return;
    for (auto* p : pyrs) p->levelInc();
    for (unsigned int i = 0; i < loop; ++i) {
      getLastFunc()();
      if (i + 1 < loop) func();
    }
    for (auto* p : pyrs) p->levelDec();
  }

 private:
  void collect(std::vector<HipaccPyramid*> acc, const std::function<void()>& func) {
    hipaccTraverse(std::move(acc), func);
  }
  template <typename... P>
  void collect(std::vector<HipaccPyramid*> acc, HipaccPyramid& p, P&&... rest) {
    acc.push_back(&p);
    collect(std::move(acc), std::forward<P>(rest)...);
  }
};

#endif  // HIPACC_HIP_TPP
