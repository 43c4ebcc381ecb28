// Spatial hash tree of finite-element surface nodes (host code).
//
// Restated from the reference's st_hash_tree (coretrace/treemesh.cpp) with
// the layout input.cpp:510-556 gives it.  The tree decides which data nodes
// are candidates for the kriging model at a ray position (the nodes of the
// zone holding the position and of its eight neighbours), so it is restated
// step for step, quirks included:
//   * zones split alternately in x and y down to the resolution limit, an
//     exhausted direction recorded as an 'x' in the zone address;
//   * neighbour addresses come from binary addition on the address digits,
//     where an out-of-range y neighbour is never detected;
//   * empty neighbour zones are created afterwards, in the iteration order of
//     a std::unordered_map keyed by address, and each receives its
//     neighbours' data through "the last zone created".
// The scene compiler flattens it for the GPU; the CPU oracle uses it as is.
#pragma once
#include <cmath>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace st_fe {

struct Zone {
    int m0 = -1, m1 = -1;        // children (-1: none); an 'x' split sets both
    bool terminal = false;
    std::vector<int> data;       // data nodes held (indices)
    std::vector<int> neighbors;  // data nodes of the surrounding zones
    std::string address;
    double xr[2] = {0, 0}, yr[2] = {0, 0};
};

struct Tree {
    double xlim[2] = {0, 0}, ylim[2] = {0, 0}, min_dx = 0, min_dy = 0;
    int nx_req = 1, ny_req = 1;
    std::vector<Zone> zones;  // zones[0] is the head; the others in creation order

    // coretrace/treemesh.cpp:353-394
    void create(double x0, double x1, double y0, double y1, double mdx, double mdy) {
        xlim[0] = x0; xlim[1] = x1; ylim[0] = y0; ylim[1] = y1;
        min_dx = mdx;
        min_dy = mdy;
        const double log2inv = 1. / log(2.);
        nx_req = (int)floor(log((xlim[1] - xlim[0]) / min_dx) * log2inv);
        nx_req = nx_req < 1 ? 1 : nx_req;
        ny_req = (int)floor(log((ylim[1] - ylim[0]) / min_dy) * log2inv);
        ny_req = ny_req < 1 ? 1 : ny_req;
        zones.assign(1, Zone());
        zones[0].xr[0] = xlim[0]; zones[0].xr[1] = xlim[1];
        zones[0].yr[0] = ylim[0]; zones[0].yr[1] = ylim[1];
    }

    int key_length() const { return (nx_req > ny_req ? nx_req : ny_req) * 2; }

    // coretrace/treemesh.cpp:757-800: halve the x and y ranges alternately,
    // '1' for the upper half
    std::string pos_to_binary(double x, double y) const {
        std::string tag;
        bool x_mode = true;
        double y0 = ylim[0], y1 = ylim[1], x0 = xlim[0], x1 = xlim[1];
        const int nc = key_length();
        for (int i = 0; i < nc; i++) {
            if (x_mode) {
                const double cx = (x0 + x1) * 0.5;
                if (x > cx) { x0 = cx; tag.append("1"); }
                else { x1 = cx; tag.append("0"); }
            } else {
                const double cy = (y0 + y1) * 0.5;
                if (y > cy) { y0 = cy; tag.append("1"); }
                else { y1 = cy; tag.append("0"); }
            }
            x_mode = !x_mode;
        }
        return tag;
    }

    // coretrace/treemesh.cpp:802-855: centre of the zone an address names
    void binary_to_pos(const std::string &b, double *x, double *y) const {
        bool x_mode = true;
        double y0 = ylim[0], y1 = ylim[1], x0 = xlim[0], x1 = xlim[1];
        *x = (x0 + x1) / 2.;
        *y = (y0 + y1) / 2.;
        for (size_t i = 0; i < b.size(); i++) {
            if (x_mode) {
                if (b[i] == '0') { x1 = *x; *x = (x1 + x0) / 2.; }
                else if (b[i] == '1') { x0 = *x; *x = (x1 + x0) / 2.; }
            } else {
                if (b[i] == '0') { y1 = *y; *y = (y1 + y0) / 2.; }
                else if (b[i] == '1') { y0 = *y; *y = (y1 + y0) / 2.; }
            }
            x_mode = !x_mode;
        }
    }

    // the zone an address leads to from the head (st_tree_node::m_proc,
    // coretrace/treemesh.cpp:209-245), -1 when a branch is missing
    int process(const std::string &key, size_t index = 0) const {
        int n = 0;
        for (;;) {
            if (index >= key.size()) return n;
            const Zone &z = zones[n];
            if (z.terminal) return n;
            const char c = key[index];
            if (c == 't') return n;
            int next;
            if (c == 'x' || c == '0') next = z.m0;
            else if (c == '1') next = z.m1;
            else return -1;
            if (next < 0) return -1;
            n = next;
            index++;
        }
    }

    // st_hash_tree::create_node (coretrace/treemesh.cpp:560-755), obj -1 for
    // an empty zone
    void create_node(int ni, size_t index, const std::string &bin, int obj, const double *dproj) {
        const double xr0 = zones[ni].xr[0], xr1 = zones[ni].xr[1], yr0 = zones[ni].yr[0], yr1 = zones[ni].yr[1];
        const double C0 = (xr0 + xr1) * 0.5, C1 = (yr0 + yr1) * 0.5;
        const double dpx = dproj ? dproj[0] : 0., dpy = dproj ? dproj[1] : 0.;
        const bool x_direction = (double)(index / 2) == ((double)index / 2.);
        int x_rec_level, y_rec_level;
        if (x_direction) {
            x_rec_level = (int)(index / 2);
            y_rec_level = x_rec_level - 1;
        } else {
            x_rec_level = y_rec_level = (int)((index - 1) / 2);
        }
        auto hold = [&](int n) {
            zones[n].terminal = true;
            if (obj >= 0) zones[n].data.push_back(obj);
        };
        if (index > bin.size() - 1) { hold(ni); return; }
        if (zones[ni].terminal) { hold(ni); return; }
        const bool up = bin[index] == '1';
        auto child = [&](int n) { return up ? zones[n].m1 : zones[n].m0; };
        // a split along the current direction, an 'x' step, or the zone holds
        auto split = [&](bool along_x) {
            const int c = child(ni);
            if (c >= 0) { create_node(c, index + 1, bin, obj, dproj); return; }
            Zone m;
            m.address = zones[ni].address + bin[index];
            m.xr[0] = xr0; m.xr[1] = xr1; m.yr[0] = yr0; m.yr[1] = yr1;
            if (along_x) { if (up) m.xr[0] = C0; else m.xr[1] = C0; }
            else { if (up) m.yr[0] = C1; else m.yr[1] = C1; }
            zones.push_back(m);
            const int mi = (int)zones.size() - 1;
            zones[ni].terminal = false;
            if (up) zones[ni].m1 = mi; else zones[ni].m0 = mi;
            create_node(mi, index + 1, bin, obj, dproj);
        };
        auto skip = [&]() {
            const int c = child(ni);
            if (c >= 0) { create_node(c, index + 1, bin, obj, dproj); return; }
            Zone m;
            m.address = zones[ni].address + "x";
            m.xr[0] = xr0; m.xr[1] = xr1; m.yr[0] = yr0; m.yr[1] = yr1;
            zones.push_back(m);
            const int mi = (int)zones.size() - 1;
            zones[ni].terminal = false;
            zones[ni].m0 = zones[ni].m1 = mi;
            create_node(mi, index + 1, bin, obj, dproj);
        };
        const bool can_x = x_rec_level < nx_req && dpx < (C0 - xr0);
        const bool can_y = y_rec_level < ny_req && dpy < (C1 - yr0);
        if (x_direction) {
            if (can_x) split(true);
            else if (can_y) skip();
            else hold(ni);
        } else {
            if (can_y) split(false);
            else if (can_x) skip();
            else hold(ni);
        }
    }

    // coretrace/treemesh.cpp:859-863
    void add_object(int obj, double x, double y, const double *objsize) {
        create_node(0, 0, pos_to_binary(x, y), obj, objsize);
    }

    // coretrace/treemesh.cpp:54-164: add (dx, dy) in {-1, 0, 1} to the x and
    // y digit strings of an address; false when x leaves the range
    static bool binary_add(const std::string &key, std::string &mod, const int val[2]) {
        int lx = 0, ly = 0;
        for (size_t i = 0; i < key.size(); i += 2) {
            if (key[i] == 'x') break;
            lx++;
        }
        for (size_t i = 1; i < key.size(); i += 2) {
            if (key[i] == 'x') break;
            ly++;
        }
        mod = key;
        int c = 0;
        bool first = true;
        for (int i = 2 * (lx - 1); i > -1; i -= 2) {
            int n = (int)(key[i] == '1');
            n += first ? val[0] : c;
            first = false;
            if (n > 1) { mod[i] = '0'; c = 1; }
            else if (n < 0) { mod[i] = '1'; c = -1; }
            else { mod[i] = n == 0 ? '0' : '1'; c = 0; break; }
            if (i == 0 && c != 0) return false;
        }
        c = 0;
        first = true;
        for (int i = 2 * (ly - 1) + 1; i > 0; i -= 2) {
            int n = (int)(key[i] == '1');
            n += first ? val[1] : c;
            first = false;
            if (n > 1) { mod[i] = '0'; c = 1; }
            else if (n < 0) { mod[i] = '1'; c = -1; }
            else { mod[i] = n == 0 ? '0' : '1'; c = 0; break; }
        }
        return true;
    }

    // coretrace/treemesh.cpp:435-509
    void add_neighborhood_data() {
        std::vector<int> tnodes;
        for (int i = 1; i < (int)zones.size(); i++)
            if (zones[i].terminal) tnodes.push_back(i);
        std::unordered_map<std::string, std::set<int>> empty_nb;
        for (int n : tnodes) {
            for (int i = -1; i < 2; i++)
                for (int j = -1; j < 2; j++) {
                    if (i == 0 && j == 0) continue;
                    const int add[2] = {i, j};
                    std::string nzone;
                    if (!binary_add(zones[n].address, nzone, add)) continue;
                    const int el = process(nzone, 0);
                    if (el >= 0) {
                        const std::vector<int> d = zones[el].data;
                        zones[n].neighbors.insert(zones[n].neighbors.end(), d.begin(), d.end());
                    } else {
                        empty_nb[nzone].insert(n);
                    }
                }
        }
        for (auto &mp : empty_nb) {
            double x, y;
            binary_to_pos(mp.first, &x, &y);
            const Zone &f = zones[*mp.second.begin()];
            const double rmin[2] = {(f.xr[1] - f.xr[0]) * 0.55, (f.yr[1] - f.yr[0]) * 0.55};
            add_object(-1, x, y, rmin);
            const int nn = (int)zones.size() - 1;
            for (int src : mp.second) {
                const std::vector<int> d = zones[src].data;
                zones[nn].neighbors.insert(zones[nn].neighbors.end(), d.begin(), d.end());
            }
        }
    }

    // coretrace/treemesh.cpp:511-537: the data of the zone holding (x, y)
    // and of its neighbours, in that order
    bool data_at(double x, double y, std::vector<int> &out) const {
        out.clear();
        const int z = process(pos_to_binary(x, y), 0);
        if (z < 0) return false;
        out.insert(out.end(), zones[z].data.begin(), zones[z].data.end());
        out.insert(out.end(), zones[z].neighbors.begin(), zones[z].neighbors.end());
        return !out.empty();
    }
};

// The layout of a node set (coretrace/input.cpp:520-556): bounds of the
// nodes, resolution 1.5 sqrt(area / n), nodes added in file order.
inline void build(Tree &t, const std::vector<double> &xyz) {
    const int n = (int)(xyz.size() / 3);
    double xmax = -1.7976931348623157e308, ymax = -1.7976931348623157e308;
    double xmin = 1.7976931348623157e308, ymin = 1.7976931348623157e308;
    for (int i = 0; i < n; i++) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1];
        xmax = x > xmax ? x : xmax;
        ymax = y > ymax ? y : ymax;
        xmin = x < xmin ? x : xmin;
        ymin = y < ymin ? y : ymin;
    }
    const double r = 1.5 * std::sqrt(((xmax - xmin) * (ymax - ymin)) / (double)n);
    t.create(xmin, xmax, ymin, ymax, r, r);
    for (int i = 0; i < n; i++) t.add_object(i, xyz[3 * i], xyz[3 * i + 1], nullptr);
    t.add_neighborhood_data();
}

}  // namespace st_fe
