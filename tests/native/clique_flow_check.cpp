// CPU check of the certificate max-flow (kclistpp_amd/csrc/clique_flow.h)
// against a plain Edmonds-Karp on the explicit network of
// kCDensestMem.c:434-467 (source -> clique (q), clique -> member (q),
// member -> sink (p)), on random graphs' k-cliques: equal flow values and
// equal source sides of the minimum cut.  Then one large instance (a
// 150-node community with edge probability 0.8, k = 5) for time and memory.
// Built and run by tests/test_host.py.
#include <cstdio>
#include <functional>
#include <cstdlib>
#include <queue>
#include <random>
#include <chrono>
#include "clique_flow.h"
using namespace std;
struct EK {
  int n; vector<vector<int>> g; vector<int> to; vector<long long> cap;
  EK(int n): n(n), g(n) {}
  void add(int a, int b, long long c) { g[a].push_back(to.size()); to.push_back(b); cap.push_back(c); g[b].push_back(to.size()); to.push_back(a); cap.push_back(0); }
  long long flow(int s, int t) { long long f = 0; for (;;) { vector<int> pe(n, -1); vector<char> vis(n, 0); queue<int> q; q.push(s); vis[s]=1;
      while (!q.empty()) { int v=q.front(); q.pop(); for (int e: g[v]) if (cap[e] && !vis[to[e]]) { vis[to[e]]=1; pe[to[e]]=e; q.push(to[e]); } }
      if (!vis[t]) return f; long long push = 1LL<<62; for (int v=t; v!=s; v=to[pe[v]^1]) push = min(push, cap[pe[v]]);
      for (int v=t; v!=s; v=to[pe[v]^1]) { cap[pe[v]]-=push; cap[pe[v]^1]+=push; } f += push; } }
  vector<char> side(int s) { vector<char> vis(n,0); queue<int> q; q.push(s); vis[s]=1; while(!q.empty()){int v=q.front();q.pop(); for(int e:g[v]) if(cap[e]&&!vis[to[e]]){vis[to[e]]=1;q.push(to[e]);}} return vis; }
};
int main() {
  mt19937_64 rng(1);
  int bad = 0;
  for (int trial = 0; trial < 400; ++trial) {
    int r = 6 + rng() % 14; int k = 2 + rng() % 4; double pe = 0.4 + (rng() % 60) / 100.0;
    vector<vector<char>> a(r, vector<char>(r, 0));
    for (int i = 0; i < r; ++i) for (int j = i+1; j < r; ++j) if ((rng() % 1000) < pe*1000) a[i][j]=a[j][i]=1;
    vector<uint32_t> mem; vector<int> cur;
    // all k-cliques
    function<void(int)> rec = [&](int start) { if ((int)cur.size()==k) { for (int x: cur) mem.push_back(x); return; }
      for (int v = start; v < r; ++v) { bool ok=true; for (int x: cur) ok = ok && a[x][v]; if (ok) { cur.push_back(v); rec(v+1); cur.pop_back(); } } };
    rec(0);
    uint64_t nc = mem.size()/k; if (!nc) continue;
    uint64_t q = 1 + rng() % r, p = 1 + rng() % (2*nc/ r + 2);
    kcl::CliqueFlow cf(k, r, nc, mem.data(), q, p);
    uint64_t f1 = cf.solve();
    EK ek(2 + r + nc);
    for (uint64_t c = 0; c < nc; ++c) { ek.add(0, 2 + r + c, q); for (int j = 0; j < k; ++j) ek.add(2 + r + c, 2 + mem[c*k+j], q); }
    for (int x = 0; x < r; ++x) ek.add(2 + x, 1, p);
    long long f2 = ek.flow(0, 1);
    auto s1 = cf.source_side(); auto s2 = ek.side(0);
    bool sideok = true; for (int x = 0; x < r; ++x) sideok = sideok && (bool)s1[x] == (bool)s2[2+x];
    if ((long long)f1 != f2 || !sideok) { ++bad; printf("trial %d r=%d k=%d nc=%lu q=%lu p=%lu: %lu vs %lld side %d\n", trial, r, k, nc, q, p, f1, f2, sideok); }
  }
  printf("mismatches: %d\n", bad);
  if (bad) return 1;
  if (getenv("CLIQUE_FLOW_SMALL_ONLY")) return 0;
  // timing: a dense community like cl1m/16's candidate (150 nodes, p = 0.8, k = 5)
  const int r = atoi(getenv("CLIQUE_FLOW_COMMUNITY") ? getenv("CLIQUE_FLOW_COMMUNITY") : "150");
  const int k = 5;
  vector<vector<char>> a(r, vector<char>(r, 0));
  for (int i = 0; i < r; ++i) for (int j = i+1; j < r; ++j) if ((rng() % 1000) < 800) a[i][j]=a[j][i]=1;
  vector<uint32_t> mem; vector<int> cur;
  function<void(int)> rec = [&](int start) { if ((int)cur.size()==k) { for (int x: cur) mem.push_back(x); return; }
    for (int v = start; v < r; ++v) { bool ok=true; for (int x: cur) ok = ok && a[x][v]; if (ok) { cur.push_back(v); rec(v+1); cur.pop_back(); } } };
  auto t0 = chrono::steady_clock::now(); rec(0); uint64_t nc = mem.size()/k;
  auto t1 = chrono::steady_clock::now();
  // candidate: the whole community, q = r, p = nc
  kcl::CliqueFlow cf(k, r, nc, mem.data(), r, nc);
  uint64_t f = cf.solve();
  auto t2 = chrono::steady_clock::now();
  auto reach = cf.reaches_sink(); int nr = 0; for (auto x: reach) nr += x;
  printf("community: %lu cliques (list %.1fs), flow %lu of %lu, saturated %d, phases %u, %.1fs, members reaching sink %d\n", nc,
         chrono::duration<double>(t1-t0).count(), f, (uint64_t)r*nc, f == (uint64_t)r*nc, cf.phases(), chrono::duration<double>(t2-t1).count(), nr);
  return 0;
}
