"""Rewrite the library's .hip sources for the CPU emulation."""
import re, sys, pathlib

def convert_launches(s):
    out, i = [], 0
    while True:
        j = s.find('<<<', i)
        if j < 0:
            out.append(s[i:]); break
        # kernel expression: identifier with optional template args, ending at j
        k = j
        depth = 0
        while k > 0:
            c = s[k - 1]
            if c == '>': depth += 1
            elif c == '<':
                if depth == 0: break
                depth -= 1
            elif depth == 0 and not (c.isalnum() or c in '_:'):
                break
            k -= 1
        kern = s[k:j]
        e = s.find('>>>', j)
        cfg = s[j + 3:e]
        assert s[e + 3] == '(', s[e:e + 20]
        p, depth = e + 4, 1
        while depth:
            if s[p] == '(': depth += 1
            elif s[p] == ')': depth -= 1
            p += 1
        args = s[e + 4:p - 1]
        parts, depth, cur = [], 0, ''
        for c in cfg:
            if c in '([{<': depth += 1
            if c in ')]}>': depth -= 1
            if c == ',' and depth == 0:
                parts.append(cur); cur = ''
            else:
                cur += c
        parts.append(cur)
        while len(parts) < 4: parts.append('0')
        out.append(s[i:k])
        kname = kern.strip()
        if args.strip():
            out.append(f'EMU_LAUNCH(({kname}), {parts[0]}, {parts[1]}, {parts[2]}, {parts[3]}, {args})')
        else:
            out.append(f'emu_launch("{kname}", dim3({parts[0]}), dim3({parts[1]}), ({parts[2]}), [=]() {{ {kname}(); }})')
        i = p
    return ''.join(out)

def convert(src, dst):
    s = pathlib.Path(src).read_text()
    s = re.sub(r'extern __shared__ (?:__align__\(\d+\) )?([\w:]+) (\w+)\[\];',
               r'\1 *\2 = static_cast<\1 *>(emu_dyn_lds());', s)
    s = convert_launches(s)
    if src.endswith('runtime.hip'):
        a = s.index('csm_status Arena::init(')
        b = s.index('// --------------------------------------------------------------- CSR build')
        s = s[:a] + ARENA + s[b:]
    pathlib.Path(dst).write_text(s)

ARENA = r'''
// emulation: one heap block per allocation, so out-of-bounds accesses are
// caught by the address sanitizer at the block's own bounds
csm_status Arena::init(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  size_ = bytes / 4096 * 4096;
  used_ = peak_ = 0;
  live_.clear();
  base_ = reinterpret_cast<char *>(1);
  return CSM_OK;
}
void Arena::release() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto &kv : live_) std::free(reinterpret_cast<void *>(kv.first));
  live_.clear();
  size_ = used_ = 0;
  base_ = nullptr;
}
void *Arena::alloc(uint64_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  bytes = (std::max<uint64_t>(bytes, 1) + 255) / 256 * 256;
  if (used_ + bytes > size_) return nullptr;
  void *p = std::malloc(bytes);
  if (!p) return nullptr;
  std::memset(p, 0xa5, bytes);
  live_[reinterpret_cast<uint64_t>(p)] = bytes;
  used_ += bytes;
  peak_ = std::max(peak_, used_);
  return p;
}
void Arena::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(reinterpret_cast<uint64_t>(p));
  if (it == live_.end()) return;
  used_ -= it->second;
  live_.erase(it);
  std::free(p);
}
uint64_t Arena::largest_free() const {
  std::lock_guard<std::mutex> g(mu_);
  return size_ - used_;
}

'''

if __name__ == '__main__':
    convert(sys.argv[1], sys.argv[2])
