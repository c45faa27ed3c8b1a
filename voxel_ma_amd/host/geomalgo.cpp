#include "geomalgo.h"

#include "../../include/vc_cc.h"
#include "vc_device.h"

#include <algorithm>
#include <set>
#include <unordered_map>

namespace util
{

void simpleTriangulate(const vector<int>& poly, vector<uTriFace>& tris)
{
    tris.clear();
    if (poly.size() < 3)
        return;
    tris.reserve(poly.size() - 2);
    int i = (int)poly.size() - 1, j = 0;
    bool inc_j = true;
    while (i - 1 > j)
    {
        if (inc_j)
        {
            tris.push_back(uTriFace(poly[i], poly[j], poly[j + 1]));
            ++j;
        }
        else
        {
            tris.push_back(uTriFace(poly[i], poly[j], poly[i - 1]));
            --i;
        }
        inc_j = !inc_j;
    }
}

void compactify(const vector<int>& v_l, vector<ivec2>& e_l,
                vector<uTriFace>& f_l)
{
    std::unordered_map<int, int> old_new;
    old_new.reserve(v_l.size());
    for (size_t n = 0; n < v_l.size(); ++n)
        old_new[v_l[n]] = (int)n;
    // unknown ids map to 0, as operator[] does in the reference
    for (auto& e : e_l)
    {
        e[0] = old_new[e[0]];
        e[1] = old_new[e[1]];
    }
    for (auto& f : f_l)
        for (int k = 0; k < 3; ++k)
            f[k] = (unsigned)old_new[(int)f[k]];
}

int findNumConnComponents(const vector<vector<int>>& adj, int n_vts,
                          vector<bool>& visited)
{
    int ncc = 0;
    vector<int> stack;
    for (int s = 0; s < n_vts; ++s)
    {
        if (visited[s])
            continue;
        ++ncc;
        stack.assign(1, s);
        while (!stack.empty())
        {
            const int u = stack.back();
            stack.pop_back();
            if (visited[u])
                continue;
            visited[u] = true;
            for (int w : adj[u])
                if (!visited[w])
                    stack.push_back(w);
        }
    }
    return ncc;
}

int findNumConnComponents(const vector<point>& vts, const vector<ivec2>& edges)
{
    vector<vector<int>> adj(vts.size());
    for (const auto& e : edges)
    {
        adj[e[0]].push_back(e[1]);
        adj[e[1]].push_back(e[0]);
    }
    vector<bool> visited(vts.size(), false);
    return findNumConnComponents(adj, (int)vts.size(), visited);
}

void computeEulerChar(const vector<uTriFace>& tris, eulerchar& ec)
{
    std::set<std::pair<int, int>> edges;
    std::set<int> vts;
    int vmax = -1;
    for (const auto& f : tris)
        for (int k = 0; k < 3; ++k)
        {
            const ivec2 e = makeEdge((int)f[k], (int)f[(k + 1) % 3]);
            edges.insert({e[0], e[1]});
            vts.insert(e[0]);
            vts.insert(e[1]);
            vmax = std::max(vmax, e[1]);
        }
    const int n = vmax + 1;
    vector<vector<int>> adj(n);
    vector<bool> visited(n, true);
    for (int v : vts)
        visited[v] = false;
    for (const auto& e : edges)
    {
        adj[e.first].push_back(e.second);
        adj[e.second].push_back(e.first);
    }
    // the reference bounds the scan by the number of referenced vertices
    // (src/geomalgo.cpp:164-165)
    const int C = findNumConnComponents(adj, (int)vts.size(), visited);
    const int V = (int)vts.size(), E = (int)edges.size(), F = (int)tris.size();
    ec = eulerchar(V, E, F, 0, C, V - E + F);
}

void computeEulerChar(const vector<point>& vts, const vector<ivec2>& e_in,
                      const vector<uTriFace>& tris, eulerchar& ec)
{
    // on the device (vc_euler_char): V = |vts|, E = distinct pairs among the
    // edges as given and the normalised triangle sides, C by union-find
    vector<int32_t> e(2 * e_in.size()), t(3 * tris.size());
    for (size_t i = 0; i < e_in.size(); ++i)
    {
        e[2 * i] = e_in[i][0];
        e[2 * i + 1] = e_in[i][1];
    }
    for (size_t i = 0; i < tris.size(); ++i)
        for (int k = 0; k < 3; ++k)
            t[3 * i + k] = (int32_t)tris[i][k];
    vc_cc_topo r{};
    vcdev::check(vc_euler_char(vcdev::context(), (int32_t)vts.size(), e.data(),
                               (int64_t)e_in.size(), t.data(), (int64_t)tris.size(), &r),
                 "vc_euler_char");
    ec = eulerchar((int)r.V, (int)r.E, (int)r.F, 0, (int)r.C, (int)r.euler);
}

} // namespace util
