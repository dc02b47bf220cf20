// Compiled-map files: the reference's `.sqwm` format, version 11 (common/bsp_map_save_load.rs:23-362,
// common/bsp_map_compact.rs:7-199, common/lightmap_compression.rs), and a Quake-format MAP writer
// (common/map_file_q1.rs:24-115) for the reference's own map_compiler.
//
// A .sqwm file is a 392-byte header (id "SqwM", version, 32 lumps of {offset, element size, element
// count}) followed by the lumps' raw little-endian #[repr(C)] arrays in lump order. Lightmaps and
// light-grid samples are stored compressed (8-bit colour components with a per-element scale).
// Files written here carry the generator's configuration as "sw_*" keys of the worldspawn entity,
// so a loaded map regenerates what the format does not hold (textures, the voxel grid the frame
// builder indexes) and takes the BSP, polygons, portals, lightmaps and light grid from the file.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "host_internal.h"

namespace
{
using namespace swh;

const char SQWM_ID[4] = {'S', 'q', 'w', 'M'};
const uint32_t SQWM_VERSION = 11;
const uint32_t MAX_LUMPS = 32;
enum
{
	LUMP_NODES = 0,
	LUMP_LEAFS = 1,
	LUMP_POLYGONS = 2,
	LUMP_PORTALS = 3,
	LUMP_LEAFS_PORTALS = 4,
	LUMP_VERTICES = 5,
	LUMP_TEXTURES = 6,
	LUMP_SUBMODELS = 7,
	LUMP_SUBMODELS_BSP_NODES = 8,
	LUMP_ENTITIES = 9,
	LUMP_KEY_VALUE_PAIRS = 10,
	LUMP_STRINGS_DATA = 11,
	LUMP_LIGHTMAPS_DATA = 12,
	LUMP_DIRECTIONAL_LIGHTMAPS_DATA = 13,
	LUMP_LIGHT_GRID_HEADER = 14,
	LUMP_LIGHT_GRID_COLUMNS = 15,
	LUMP_LIGHT_GRID_SAMPLES = 16,
	NUM_USED_LUMPS = 17,
};

struct Lump
{
	uint32_t offset, element_size, element_count;
};
struct Header
{
	char id[4];
	uint32_t version;
	Lump lumps[MAX_LUMPS];
};
static_assert(sizeof(Header) == 392, "BspMapHeader layout");

// The #[repr(C)] element types of bsp_map_compact.rs that the front-end does not already mirror.
struct Submodel
{
	uint32_t root_node, first_polygon, num_polygons;
};
struct SubmodelBSPNode
{
	Plane plane;
	uint32_t first_polygon, num_polygons;
	uint32_t children[2];
};
struct Entity
{
	uint32_t first_key_value_pair, num_key_value_pairs, submodel_index;
};
struct StringRef
{
	uint16_t offset, size;
};
struct KeyValuePair
{
	StringRef key, value;
};
const uint32_t TEXTURE_NAME_LEN = 64;
struct LightGridHeader
{
	float grid_cell_size[3];
	float grid_start[3];
	uint32_t grid_size[3];
};
struct LightGridColumn
{
	uint32_t start_z, first_sample, num_samples;
};

// lightmap_compression.rs
struct CompressedColor
{
	uint8_t components[3];
	uint8_t scale;
};
struct CompressedVector
{
	int8_t components[3];
	uint8_t scale;
};
struct DirectionalLightmapElementCompressed
{
	CompressedColor ambient_light;
	CompressedVector light_direction_vector_scaled;
	uint8_t directional_light_deviation;
	CompressedColor directional_light_color;
};
struct LightGridElementCompressed
{
	CompressedColor light_cube[6];
	CompressedVector light_direction_vector_scaled;
	CompressedColor directional_light_color;
};
static_assert(sizeof(Node) == 24 && sizeof(Leaf) == 16 && sizeof(Polygon) == 80 && sizeof(Portal) == 32 && sizeof(V3) == 12,
	"compact map element layouts");
static_assert(sizeof(DirectionalLightmapElementCompressed) == 13 && sizeof(LightGridElementCompressed) == 32, "compressed layouts");

const float COLOR_SCALE = 255.0f, VECTOR_SCALE = 127.0f, LIGHT_DEVIATION_SCALE = 255.0f;

// Rust `x as u8` / `x as i8` from f32: truncation toward zero, saturating, NaN -> 0.
uint8_t f32_as_u8(float x) { return !(x > 0.0f) ? 0 : (x >= 255.0f ? 255 : uint8_t(x)); }
int8_t f32_as_i8(float x) { return x != x ? 0 : (x <= -128.0f ? -128 : (x >= 127.0f ? 127 : int8_t(x))); }

// CompressedColor::compress / decompress
CompressedColor compress_color(const float c[3])
{
	const float max_component = std::fmax(std::fmax(c[0], c[1]), c[2]);
	const float scale = std::fmin(std::fmax(std::ceil(max_component), 1.0f), 255.0f);
	const float multiplier = COLOR_SCALE / scale;
	CompressedColor r;
	for(int i = 0; i < 3; ++i)
		r.components[i] = f32_as_u8(std::fmin(std::fmax(c[i] * multiplier, 0.0f), 255.0f));
	r.scale = f32_as_u8(scale);
	return r;
}
void decompress_color(const CompressedColor& c, float out[3])
{
	const float multiplier = float(c.scale) / COLOR_SCALE;
	for(int i = 0; i < 3; ++i)
		out[i] = float(c.components[i]) * multiplier;
}
// CompressedVector::compress / decompress
CompressedVector compress_vector(V3 v)
{
	const float max_component = std::fmax(std::fmax(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z));
	const float scale = std::fmin(std::fmax(std::ceil(max_component), 1.0f), 255.0f);
	const float multiplier = VECTOR_SCALE / scale;
	const float c[3] = {v.x, v.y, v.z};
	CompressedVector r;
	for(int i = 0; i < 3; ++i)
		r.components[i] = f32_as_i8(std::fmin(std::fmax(c[i] * multiplier, -127.0f), 127.0f));
	r.scale = f32_as_u8(scale);
	return r;
}
V3 decompress_vector(const CompressedVector& c)
{
	const float multiplier = float(c.scale) / VECTOR_SCALE;
	return v3(float(c.components[0]) * multiplier, float(c.components[1]) * multiplier, float(c.components[2]) * multiplier);
}

DirectionalLightmapElementCompressed compress_dir(const sw_directional_lightmap_element& e)
{
	DirectionalLightmapElementCompressed r;
	r.ambient_light = compress_color(e.ambient_light);
	r.light_direction_vector_scaled =
		compress_vector(v3(e.light_direction_vector_scaled[0], e.light_direction_vector_scaled[1], e.light_direction_vector_scaled[2]));
	r.directional_light_deviation = f32_as_u8(std::fmin(std::fmax(e.directional_light_deviation * LIGHT_DEVIATION_SCALE, 0.0f), 255.0f));
	r.directional_light_color = compress_color(e.directional_light_color);
	return r;
}
sw_directional_lightmap_element decompress_dir(const DirectionalLightmapElementCompressed& c)
{
	sw_directional_lightmap_element e;
	decompress_color(c.ambient_light, e.ambient_light);
	V3 d = decompress_vector(c.light_direction_vector_scaled);
	// the reference's fix-up: the vector must not be zero after quantisation
	if(dot(d, d) <= 0.0f)
		d = v3(0.0f, 0.0f, 1.0f / 1024.0f);
	e.light_direction_vector_scaled[0] = d.x;
	e.light_direction_vector_scaled[1] = d.y;
	e.light_direction_vector_scaled[2] = d.z;
	e.directional_light_deviation = float(c.directional_light_deviation) / LIGHT_DEVIATION_SCALE;
	decompress_color(c.directional_light_color, e.directional_light_color);
	return e;
}
LightGridElementCompressed compress_grid(const LightGridElement& e)
{
	LightGridElementCompressed r;
	for(int s = 0; s < 6; ++s)
		r.light_cube[s] = compress_color(e.light_cube[s]);
	r.light_direction_vector_scaled = compress_vector(e.light_direction_vector_scaled);
	r.directional_light_color = compress_color(e.directional_light_color);
	return r;
}
LightGridElement decompress_grid(const LightGridElementCompressed& c)
{
	LightGridElement e;
	for(int s = 0; s < 6; ++s)
		decompress_color(c.light_cube[s], e.light_cube[s]);
	e.light_direction_vector_scaled = decompress_vector(c.light_direction_vector_scaled);
	decompress_color(c.directional_light_color, e.directional_light_color);
	return e;
}

// ---------------------------------------------------------------------------------------------
// writing
// ---------------------------------------------------------------------------------------------
const float CELL = 128.0f;

std::string fmt_float(float x)
{
	char buf[64];
	std::snprintf(buf, sizeof(buf), "%.9g", double(x));
	return buf;
}

// The generator configuration as worldspawn keys (round-trips exactly: floats in 9 digits).
std::vector<std::pair<std::string, std::string>> generator_keys(const swh_map_config& c)
{
	auto u = [](uint32_t v) { return std::to_string(v); };
	return {{"sw_seed", u(c.seed)}, {"sw_grid", u(c.grid[0]) + " " + u(c.grid[1]) + " " + u(c.grid[2])},
		{"sw_solid_probability", fmt_float(c.solid_probability)}, {"sw_num_lights", u(c.num_lights)},
		{"sw_light_intensity", fmt_float(c.light_intensity_min) + " " + fmt_float(c.light_intensity_max)},
		{"sw_sun", u(c.sun)}, {"sw_sky_fraction", fmt_float(c.sky_fraction)}, {"sw_texture_size", u(c.texture_size)},
		{"sw_num_materials", u(c.num_materials)}, {"sw_normal_maps", u(c.normal_maps)},
		{"sw_specular_fraction", fmt_float(c.specular_fraction)}, {"sw_alpha_blend_fraction", fmt_float(c.alpha_blend_fraction)},
		{"sw_directional_lightmaps", u(c.directional_lightmaps)}, {"sw_num_decal_textures", u(c.num_decal_textures)},
		{"sw_num_model_textures", u(c.num_model_textures)}, {"sw_num_submodels", u(c.num_submodels)}};
}

std::string material_name(const swh_map& m, uint32_t i)
{
	return m.materials[i].sky ? std::string("sw_gen/sky") : "sw_gen/material_" + std::to_string(i);
}

struct Writer
{
	std::vector<uint8_t> bytes;
	Header header{};
	Writer()
	{
		bytes.resize(sizeof(Header));
		std::memcpy(header.id, SQWM_ID, 4);
		header.version = SQWM_VERSION;
	}
	// write_lump (bsp_map_save_load.rs:291-317): offset, element size, count, then the raw array
	template <class T>
	void lump(uint32_t index, const T* data, size_t count)
	{
		Lump& l = header.lumps[index];
		l.offset = uint32_t(bytes.size());
		l.element_size = uint32_t(sizeof(T));
		l.element_count = uint32_t(count);
		const uint8_t* p = reinterpret_cast<const uint8_t*>(data);
		bytes.insert(bytes.end(), p, p + sizeof(T) * count);
	}
	template <class T>
	void lump(uint32_t index, const std::vector<T>& v)
	{
		lump(index, v.data(), v.size());
	}
	void finish() { std::memcpy(bytes.data(), &header, sizeof(Header)); }
};

struct Strings
{
	std::vector<uint8_t> data;
	std::map<std::string, StringRef> seen;
	bool overflow = false;
	StringRef add(const std::string& s)
	{
		auto it = seen.find(s);
		if(it != seen.end())
			return it->second;
		if(data.size() + s.size() > 65535 || s.size() > 65535)
		{
			overflow = true;
			return StringRef{0, 0};
		}
		const StringRef r{uint16_t(data.size()), uint16_t(s.size())};
		data.insert(data.end(), s.begin(), s.end());
		seen[s] = r;
		return r;
	}
};

int build_sqwm(const swh_map& m, std::vector<uint8_t>& out)
{
	Writer w;
	w.lump(LUMP_NODES, m.nodes);
	w.lump(LUMP_LEAFS, m.leafs);
	w.lump(LUMP_POLYGONS, m.polygons);
	w.lump(LUMP_PORTALS, m.portals);
	w.lump(LUMP_LEAFS_PORTALS, m.leafs_portals);
	w.lump(LUMP_VERTICES, m.vertices);
	std::vector<std::array<char, TEXTURE_NAME_LEN>> textures(m.materials.size());
	for(size_t i = 0; i < m.materials.size(); ++i)
	{
		textures[i].fill(0);
		const std::string name = material_name(m, uint32_t(i));
		std::memcpy(textures[i].data(), name.data(), std::min<size_t>(name.size(), TEXTURE_NAME_LEN));
	}
	w.lump(LUMP_TEXTURES, textures);
	{
		std::vector<Submodel> sub(m.submodels.size());
		for(size_t i = 0; i < sub.size(); ++i)
			sub[i] = Submodel{m.submodels[i].root_node, m.submodels[i].first_polygon, m.submodels[i].num_polygons};
		std::vector<SubmodelBSPNode> nodes(m.submodel_nodes.size());
		for(size_t i = 0; i < nodes.size(); ++i)
		{
			const swh::SubmodelNode& n = m.submodel_nodes[i];
			nodes[i] = SubmodelBSPNode{n.plane, n.first_polygon, n.num_polygons, {n.children[0], n.children[1]}};
		}
		w.lump(LUMP_SUBMODELS, sub);
		w.lump(LUMP_SUBMODELS_BSP_NODES, nodes);
	}

	// entities: worldspawn (with the generator's keys), then one "light" per static point light
	Strings strings;
	std::vector<Entity> entities;
	std::vector<KeyValuePair> kvs;
	auto entity = [&](const std::vector<std::pair<std::string, std::string>>& keys, uint32_t submodel = 0xFFFFFFFFu) {
		entities.push_back(Entity{uint32_t(kvs.size()), uint32_t(keys.size()), submodel});
		for(const auto& kv : keys)
			kvs.push_back(KeyValuePair{strings.add(kv.first), strings.add(kv.second)});
	};
	auto world = generator_keys(m.config);
	world.insert(world.begin(), {"classname", "worldspawn"});
	entity(world);
	for(const PointLight& l : m.lights)
		entity({{"classname", "light"}, {"origin", fmt_float(l.pos.x) + " " + fmt_float(l.pos.y) + " " + fmt_float(l.pos.z)},
			{"light", fmt_float(l.intensity / 32.0f)}, {"color", "255 255 255"}});
	// one brush entity per inline model (doors and platforms alternate), holding its submodel index
	for(uint32_t k = 0; k < m.submodels.size(); ++k)
		entity({{"classname", k % 2 == 0 ? "func_door" : "func_plat"}}, k);
	if(strings.overflow)
		return -1;
	w.lump(LUMP_ENTITIES, entities);
	w.lump(LUMP_KEY_VALUE_PAIRS, kvs);
	w.lump(LUMP_STRINGS_DATA, strings.data);

	std::vector<CompressedColor> lm(m.lm_simple.size() / 3);
	for(size_t i = 0; i < lm.size(); ++i)
		lm[i] = compress_color(&m.lm_simple[i * 3]);
	w.lump(LUMP_LIGHTMAPS_DATA, lm);
	std::vector<DirectionalLightmapElementCompressed> dlm(m.lm_dir.size());
	for(size_t i = 0; i < dlm.size(); ++i)
		dlm[i] = compress_dir(m.lm_dir[i]);
	w.lump(LUMP_DIRECTIONAL_LIGHTMAPS_DATA, dlm);

	// the light grid: one sample per cell centre, stored per (x, y) column from the lowest to the
	// highest non-zero sample (solid cells hold zero samples, which the reader skips)
	LightGridHeader gh{};
	std::vector<LightGridColumn> columns;
	std::vector<LightGridElementCompressed> samples;
	if(!m.light_grid.empty())
	{
		for(int k = 0; k < 3; ++k)
		{
			gh.grid_cell_size[k] = CELL;
			gh.grid_start[k] = CELL * 0.5f;
		}
		gh.grid_size[0] = m.gx;
		gh.grid_size[1] = m.gy;
		gh.grid_size[2] = m.gz;
		columns.resize(size_t(m.gx) * m.gy);
		for(uint32_t y = 0; y < m.gy; ++y)
			for(uint32_t x = 0; x < m.gx; ++x)
			{
				int32_t z0 = -1, z1 = -1;
				for(uint32_t z = 0; z < m.gz; ++z)
				{
					const LightGridElement& e = m.light_grid[cell_index(m, x, y, z)];
					const float* p = &e.light_cube[0][0];
					bool nonzero = false;
					for(int i = 0; i < 18 && !nonzero; ++i)
						nonzero = p[i] != 0.0f;
					if(nonzero)
					{
						if(z0 < 0)
							z0 = int32_t(z);
						z1 = int32_t(z);
					}
				}
				LightGridColumn& c = columns[size_t(y) * m.gx + x];
				c.first_sample = uint32_t(samples.size());
				c.start_z = z0 < 0 ? 0 : uint32_t(z0);
				c.num_samples = z0 < 0 ? 0 : uint32_t(z1 - z0 + 1);
				for(uint32_t z = c.start_z; z < c.start_z + c.num_samples; ++z)
					samples.push_back(compress_grid(m.light_grid[cell_index(m, x, y, z)]));
			}
	}
	w.lump(LUMP_LIGHT_GRID_HEADER, &gh, 1);
	w.lump(LUMP_LIGHT_GRID_COLUMNS, columns);
	w.lump(LUMP_LIGHT_GRID_SAMPLES, samples);
	w.finish();
	out.swap(w.bytes);
	return 0;
}

// ---------------------------------------------------------------------------------------------
// reading
// ---------------------------------------------------------------------------------------------
struct Reader
{
	const std::vector<uint8_t>& bytes;
	Header header{};
	std::string error;
	// read_lump (bsp_map_save_load.rs:319-347), with the element size and the range checked
	template <class T>
	bool lump(uint32_t index, std::vector<T>& out)
	{
		const Lump& l = header.lumps[index];
		if(l.element_count == 0)
		{
			out.clear();
			return true;
		}
		if(l.element_size != sizeof(T))
		{
			error = "lump " + std::to_string(index) + ": element size " + std::to_string(l.element_size) + ", expected " +
				std::to_string(sizeof(T));
			return false;
		}
		if(uint64_t(l.offset) + uint64_t(l.element_size) * l.element_count > bytes.size())
		{
			error = "lump " + std::to_string(index) + " beyond the end of the file";
			return false;
		}
		out.resize(l.element_count);
		std::memcpy(out.data(), bytes.data() + l.offset, sizeof(T) * out.size());
		return true;
	}
};

std::string get_string(const std::vector<uint8_t>& data, StringRef r)
{
	if(uint32_t(r.offset) + r.size > data.size())
		return std::string();
	return std::string(reinterpret_cast<const char*>(data.data()) + r.offset, r.size);
}

bool parse_u32s(const std::string& s, uint32_t* out, int n)
{
	const char* p = s.c_str();
	for(int i = 0; i < n; ++i)
	{
		char* end = nullptr;
		const unsigned long v = std::strtoul(p, &end, 10);
		if(end == p)
			return false;
		out[i] = uint32_t(v);
		p = end;
	}
	return true;
}
bool parse_floats(const std::string& s, float* out, int n)
{
	const char* p = s.c_str();
	for(int i = 0; i < n; ++i)
	{
		char* end = nullptr;
		const float v = std::strtof(p, &end);
		if(end == p)
			return false;
		out[i] = v;
		p = end;
	}
	return true;
}

// Checks every index a frame would follow: node children, leaf polygon / portal ranges, polygon
// vertex ranges and textures, portal leafs and vertices, lightmap offsets.
bool validate(const swh_map& m, std::string& error)
{
	const uint32_t nn = uint32_t(m.nodes.size()), nl = uint32_t(m.leafs.size()), np = uint32_t(m.polygons.size());
	if(nn == 0 || nl == 0)
		return (error = "map without nodes or leafs"), false;
	for(const Node& n : m.nodes)
		for(uint32_t c : n.children)
			if(c >= FIRST_LEAF_INDEX ? c - FIRST_LEAF_INDEX >= nl : c >= nn)
				return (error = "node child out of range"), false;
	// The walks descend from the root (the last node, bsp_map_compact.rs:289-292) without a depth
	// limit: every node and leaf must be reached exactly once, so a child pointing back at an ancestor
	// (a cycle) or a shared subtree is an error.
	{
		std::vector<uint8_t> seen_node(nn, 0), seen_leaf(nl, 0);
		std::vector<uint32_t> stack{nn - 1};
		while(!stack.empty())
		{
			const uint32_t index = stack.back();
			stack.pop_back();
			if(index >= FIRST_LEAF_INDEX)
			{
				if(seen_leaf[index - FIRST_LEAF_INDEX]++)
					return (error = "BSP leaf reached twice: the nodes do not form a tree"), false;
				continue;
			}
			if(seen_node[index]++)
				return (error = "BSP node reached twice: the nodes do not form a tree"), false;
			stack.push_back(m.nodes[index].children[0]);
			stack.push_back(m.nodes[index].children[1]);
		}
	}
	for(const Leaf& l : m.leafs)
		if(uint64_t(l.first_polygon) + l.num_polygons > np || uint64_t(l.first_leaf_portal) + l.num_leaf_portals > m.leafs_portals.size())
			return (error = "leaf polygon or portal range out of range"), false;
	for(uint32_t p : m.leafs_portals)
		if(p >= m.portals.size())
			return (error = "leaf portal index out of range"), false;
	for(const Polygon& p : m.polygons)
	{
		if(p.num_vertices < 3 || p.num_vertices > SW_MAX_POLYGON_VERTICES || uint64_t(p.first_vertex) + p.num_vertices > m.vertices.size())
			return (error = "polygon vertex range out of range"), false;
		if(p.texture >= m.materials.size())
			return (error = "polygon texture out of range"), false;
		if(p.lightmap_data_offset != 0)
		{
			// get_lightmap_size (lightmap.rs:19-27) of the polygon's texture-coordinate extent: the
			// surface builder and the decal / model light lookups read lw * lh elements from the offset
			const int64_t lw = ((int64_t(p.tex_coord_max[0]) + LIGHTMAP_SCALE - 1) >> LIGHTMAP_SCALE_LOG2) -
				(int64_t(p.tex_coord_min[0]) >> LIGHTMAP_SCALE_LOG2) + 1;
			const int64_t lh = ((int64_t(p.tex_coord_max[1]) + LIGHTMAP_SCALE - 1) >> LIGHTMAP_SCALE_LOG2) -
				(int64_t(p.tex_coord_min[1]) >> LIGHTMAP_SCALE_LOG2) + 1;
			const uint64_t count = m.lm_dir.empty() ? m.lm_simple.size() / 3 : m.lm_dir.size();
			if(lw < 1 || lh < 1 || lw > SW_MAX_LIGHTMAP_SIZE + 2 || lh > SW_MAX_LIGHTMAP_SIZE + 2)
				return (error = "polygon lightmap size out of range"), false;
			if(uint64_t(p.lightmap_data_offset) + uint64_t(lw * lh) > count)
				return (error = "polygon lightmap extends past the lightmap data"), false;
		}
	}
	for(const Portal& p : m.portals)
		if(p.leafs[0] >= nl || p.leafs[1] >= nl || uint64_t(p.first_vertex) + p.num_vertices > m.vertices.size())
			return (error = "portal out of range"), false;
	// inline models: polygon ranges, and node trees walked from each root reaching every node once
	const uint32_t nsn = uint32_t(m.submodel_nodes.size());
	for(const swh::SubmodelNode& n : m.submodel_nodes)
		if(uint64_t(n.first_polygon) + n.num_polygons > np)
			return (error = "submodel node polygon range out of range"), false;
	std::vector<uint8_t> seen_sub(nsn, 0);
	for(const swh::Submodel& sm : m.submodels)
	{
		if(uint64_t(sm.first_polygon) + sm.num_polygons > np || sm.root_node >= nsn)
			return (error = "submodel out of range"), false;
		std::vector<uint32_t> stack{sm.root_node};
		while(!stack.empty())
		{
			const uint32_t i = stack.back();
			stack.pop_back();
			if(seen_sub[i]++)
				return (error = "submodel node reached twice: the submodel nodes do not form trees"), false;
			for(uint32_t c : m.submodel_nodes[i].children)
				if(c < nsn)
					stack.push_back(c);
		}
	}
	return true;
}

int load_sqwm(const std::vector<uint8_t>& bytes, swh_map** out, std::string& error)
{
	Reader r{bytes};
	if(bytes.size() < sizeof(Header))
		return (error = "file shorter than the map header"), -1;
	std::memcpy(&r.header, bytes.data(), sizeof(Header));
	if(std::memcmp(r.header.id, SQWM_ID, 4) != 0)
		return (error = "not a Square-Wheel map (id)"), -1;
	if(r.header.version != SQWM_VERSION)
		return (error = "map version " + std::to_string(r.header.version) + ", expected 11"), -1;

	std::vector<Node> nodes;
	std::vector<Leaf> leafs;
	std::vector<Polygon> polygons;
	std::vector<Portal> portals;
	std::vector<uint32_t> leafs_portals;
	std::vector<V3> vertices;
	std::vector<std::array<char, TEXTURE_NAME_LEN>> textures;
	std::vector<Submodel> submodels;
	std::vector<SubmodelBSPNode> submodel_nodes;
	std::vector<Entity> entities;
	std::vector<KeyValuePair> kvs;
	std::vector<uint8_t> strings;
	std::vector<CompressedColor> lm;
	std::vector<DirectionalLightmapElementCompressed> dlm;
	std::vector<LightGridHeader> gh;
	std::vector<LightGridColumn> columns;
	std::vector<LightGridElementCompressed> samples;
	if(!r.lump(LUMP_NODES, nodes) || !r.lump(LUMP_LEAFS, leafs) || !r.lump(LUMP_POLYGONS, polygons) || !r.lump(LUMP_PORTALS, portals) ||
		!r.lump(LUMP_LEAFS_PORTALS, leafs_portals) || !r.lump(LUMP_VERTICES, vertices) || !r.lump(LUMP_TEXTURES, textures) ||
		!r.lump(LUMP_SUBMODELS, submodels) || !r.lump(LUMP_SUBMODELS_BSP_NODES, submodel_nodes) || !r.lump(LUMP_ENTITIES, entities) || !r.lump(LUMP_KEY_VALUE_PAIRS, kvs) || !r.lump(LUMP_STRINGS_DATA, strings) ||
		!r.lump(LUMP_LIGHTMAPS_DATA, lm) || !r.lump(LUMP_DIRECTIONAL_LIGHTMAPS_DATA, dlm) || !r.lump(LUMP_LIGHT_GRID_HEADER, gh) ||
		!r.lump(LUMP_LIGHT_GRID_COLUMNS, columns) || !r.lump(LUMP_LIGHT_GRID_SAMPLES, samples))
		return (error = r.error), -1;

	// the worldspawn's generator keys
	std::map<std::string, std::string> world;
	for(const Entity& e : entities)
	{
		if(uint64_t(e.first_key_value_pair) + e.num_key_value_pairs > kvs.size())
			return (error = "entity key range out of range"), -1;
		std::map<std::string, std::string> keys;
		for(uint32_t k = 0; k < e.num_key_value_pairs; ++k)
			keys[get_string(strings, kvs[e.first_key_value_pair + k].key)] = get_string(strings, kvs[e.first_key_value_pair + k].value);
		if(keys["classname"] == "worldspawn")
			world = keys;
	}
	swh_map_config c{};
	float intensity[2];
	if(!world.count("sw_seed") || !parse_u32s(world["sw_seed"], &c.seed, 1) || !parse_u32s(world["sw_grid"], c.grid, 3) ||
		!parse_floats(world["sw_solid_probability"], &c.solid_probability, 1) || !parse_u32s(world["sw_num_lights"], &c.num_lights, 1) ||
		!parse_floats(world["sw_light_intensity"], intensity, 2) || !parse_u32s(world["sw_sun"], &c.sun, 1) ||
		!parse_floats(world["sw_sky_fraction"], &c.sky_fraction, 1) || !parse_u32s(world["sw_texture_size"], &c.texture_size, 1) ||
		!parse_u32s(world["sw_num_materials"], &c.num_materials, 1) || !parse_u32s(world["sw_normal_maps"], &c.normal_maps, 1) ||
		!parse_floats(world["sw_specular_fraction"], &c.specular_fraction, 1) ||
		!parse_floats(world["sw_alpha_blend_fraction"], &c.alpha_blend_fraction, 1) ||
		!parse_u32s(world["sw_directional_lightmaps"], &c.directional_lightmaps, 1) ||
		!parse_u32s(world["sw_num_decal_textures"], &c.num_decal_textures, 1) ||
		!parse_u32s(world["sw_num_model_textures"], &c.num_model_textures, 1))
		return (error = "the worldspawn entity has no generator configuration (sw_* keys): textures cannot be rebuilt"), -1;
	if(world.count("sw_num_submodels") && !parse_u32s(world["sw_num_submodels"], &c.num_submodels, 1))
		return (error = "bad sw_num_submodels key"), -1;
	c.light_intensity_min = intensity[0];
	c.light_intensity_max = intensity[1];

	swh_map* m = nullptr;
	if(swh_map_create(&c, &m) != 0 || !m)
		return (error = "the generator configuration is invalid"), -1;
	if(nodes.size() != m->nodes.size() || leafs.size() != m->leafs.size() || polygons.size() != m->polygons.size() ||
		portals.size() != m->portals.size() || textures.size() != m->materials.size() || lm.size() != m->lm_simple.size() / 3 ||
		dlm.size() != m->lm_dir.size() || (!gh.empty() && (gh[0].grid_size[0] != m->gx || gh[0].grid_size[1] != m->gy || gh[0].grid_size[2] != m->gz)))
	{
		swh_map_destroy(m);
		return (error = "the map's lumps do not match its generator configuration"), -1;
	}
	// texture names -> materials
	std::vector<uint32_t> tex_material(textures.size());
	for(size_t i = 0; i < textures.size(); ++i)
	{
		const std::string name(textures[i].data(), strnlen(textures[i].data(), TEXTURE_NAME_LEN));
		uint32_t mi = 0xFFFFFFFFu;
		for(uint32_t k = 0; k < m->materials.size(); ++k)
			if(material_name(*m, k) == name)
				mi = k;
		if(mi == 0xFFFFFFFFu)
		{
			swh_map_destroy(m);
			return (error = "unknown texture name \"" + name + "\""), -1;
		}
		tex_material[i] = mi;
	}
	for(Polygon& p : polygons)
		if(p.texture < tex_material.size())
			p.texture = tex_material[p.texture];
		else
			p.texture = 0xFFFFFFFFu; // rejected by the validation below

	// the file's data replace the generated ones
	m->nodes.swap(nodes);
	m->leafs.swap(leafs);
	m->polygons.swap(polygons);
	m->portals.swap(portals);
	m->leafs_portals.swap(leafs_portals);
	m->vertices.swap(vertices);
	m->submodels.resize(submodels.size());
	for(size_t i = 0; i < submodels.size(); ++i)
		m->submodels[i] = swh::Submodel{submodels[i].root_node, submodels[i].first_polygon, submodels[i].num_polygons};
	m->submodel_nodes.resize(submodel_nodes.size());
	for(size_t i = 0; i < submodel_nodes.size(); ++i)
	{
		const SubmodelBSPNode& n = submodel_nodes[i];
		m->submodel_nodes[i] = swh::SubmodelNode{n.plane, n.first_polygon, n.num_polygons, {n.children[0], n.children[1]}};
	}
	// the polygons' parents from the file: leaf ranges, the rest belongs to inline models
	m->polygon_leaf.assign(m->polygons.size(), NO_POLYGON_LEAF);
	for(uint32_t l = 0; l < m->leafs.size(); ++l)
		for(uint32_t k = 0; k < m->leafs[l].num_polygons && uint64_t(m->leafs[l].first_polygon) + k < m->polygons.size(); ++k)
			m->polygon_leaf[m->leafs[l].first_polygon + k] = l;
	for(size_t i = 0; i < lm.size(); ++i)
		decompress_color(lm[i], &m->lm_simple[i * 3]);
	for(size_t i = 0; i < dlm.size(); ++i)
		m->lm_dir[i] = decompress_dir(dlm[i]);
	m->light_grid.assign(m->light_grid.size(), LightGridElement{});
	if(!gh.empty())
	{
		if(columns.size() != size_t(m->gx) * m->gy)
		{
			swh_map_destroy(m);
			return (error = "light grid columns do not match the grid size"), -1;
		}
		for(uint32_t y = 0; y < m->gy; ++y)
			for(uint32_t x = 0; x < m->gx; ++x)
			{
				const LightGridColumn& col = columns[size_t(y) * m->gx + x];
				if(uint64_t(col.start_z) + col.num_samples > m->gz || uint64_t(col.first_sample) + col.num_samples > samples.size())
				{
					swh_map_destroy(m);
					return (error = "light grid column out of range"), -1;
				}
				for(uint32_t k = 0; k < col.num_samples; ++k)
					m->light_grid[cell_index(*m, x, y, col.start_z + k)] = decompress_grid(samples[col.first_sample + k]);
			}
	}
	if(!validate(*m, error))
	{
		swh_map_destroy(m);
		return -1;
	}
	*out = m;
	return 0;
}

// ---------------------------------------------------------------------------------------------
// Quake MAP
// ---------------------------------------------------------------------------------------------
// One brush plane through three points whose (p0 - p1) x (p2 - p1) is the outward normal
// (map_polygonizer.rs:293-305: a point is inside the brush where dot(p, vec) <= dist).
void brush_plane(std::string& s, const int p[3], int axis, int sign, int len, const std::string& texture)
{
	const int a = (axis + 1) % 3, b = (axis + 2) % 3; // a x b = +axis
	int p1[3] = {p[0], p[1], p[2]}, p0[3] = {p[0], p[1], p[2]}, p2[3] = {p[0], p[1], p[2]};
	if(sign > 0)
	{
		p0[a] += len;
		p2[b] += len;
	}
	else
	{
		p0[b] += len;
		p2[a] += len;
	}
	char buf[256];
	std::snprintf(buf, sizeof(buf), "( %d %d %d ) ( %d %d %d ) ( %d %d %d ) %s 0 0 0 1 1\n", p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], p2[0],
		p2[1], p2[2], texture.c_str());
	s += buf;
}

std::string build_quake_map(const swh_map& m)
{
	std::string s = "// Square-Wheel voxel-grid map, seed " + std::to_string(m.config.seed) + "\n{\n\"classname\" \"worldspawn\"\n";
	for(const auto& kv : generator_keys(m.config))
		s += "\"" + kv.first + "\" \"" + kv.second + "\"\n";
	const int C = int(CELL);
	const uint32_t nmat = uint32_t(m.materials.size()) - 1;
	for(uint32_t z = 0; z < m.gz; ++z)
		for(uint32_t y = 0; y < m.gy; ++y)
			for(uint32_t x0 = 0; x0 < m.gx;)
			{
				const uint8_t kind = m.cell[cell_index(m, x0, y, z)];
				uint32_t x1 = x0 + 1;
				while(x1 < m.gx && m.cell[cell_index(m, x1, y, z)] == kind)
					++x1;
				if(kind != 0)
				{
					// one box brush per run of solid (or sky) cells along x
					const std::string tex =
						kind == 2 ? std::string("sw_gen/sky") : "sw_gen/material_" + std::to_string((x0 * 73856093u ^ y * 19349663u ^ z * 83492791u) % nmat);
					const int lo[3] = {int(x0) * C, int(y) * C, int(z) * C}, hi[3] = {int(x1) * C, int(y + 1) * C, int(z + 1) * C};
					s += "{\n";
					for(int axis = 0; axis < 3; ++axis)
					{
						brush_plane(s, lo, axis, -1, C, tex);
						brush_plane(s, hi, axis, +1, C, tex);
					}
					s += "}\n";
				}
				x0 = x1;
			}
	s += "}\n";
	for(const PointLight& l : m.lights)
		s += "{\n\"classname\" \"light\"\n\"origin\" \"" + fmt_float(l.pos.x) + " " + fmt_float(l.pos.y) + " " + fmt_float(l.pos.z) +
			"\"\n\"light\" \"" + fmt_float(l.intensity / 32.0f) + "\"\n\"color\" \"255 255 255\"\n}\n";
	// inline models as brush entities: one box brush each, its faces' textures
	for(uint32_t k = 0; k < m.submodels.size(); ++k)
	{
		const swh::Submodel& sm = m.submodels[k];
		int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
		for(uint32_t pi = sm.first_polygon; pi < sm.first_polygon + sm.num_polygons; ++pi)
			for(uint32_t v = 0; v < m.polygons[pi].num_vertices; ++v)
			{
				const V3 q = m.vertices[m.polygons[pi].first_vertex + v];
				const int c3[3] = {int(std::lround(q.x)), int(std::lround(q.y)), int(std::lround(q.z))};
				for(int a = 0; a < 3; ++a)
				{
					lo[a] = std::min(lo[a], c3[a]);
					hi[a] = std::max(hi[a], c3[a]);
				}
			}
		const std::string tex = material_name(m, m.polygons[sm.first_polygon].texture);
		s += std::string("{\n\"classname\" \"") + (k % 2 == 0 ? "func_door" : "func_plat") + "\"\n{\n";
		for(int axis = 0; axis < 3; ++axis)
		{
			brush_plane(s, lo, axis, -1, 16, tex);
			brush_plane(s, hi, axis, +1, 16, tex);
		}
		s += "}\n}\n";
	}
	if(!m.empty_cells.empty())
	{
		const uint32_t c = m.empty_cells[0];
		const uint32_t x = c % m.gx, y = (c / m.gx) % m.gy, z = c / (m.gx * m.gy);
		s += "{\n\"classname\" \"info_player_start\"\n\"origin\" \"" + std::to_string(int(x) * C + C / 2) + " " +
			std::to_string(int(y) * C + C / 2) + " " + std::to_string(int(z) * C + C / 2) + "\"\n}\n";
	}
	return s;
}

bool write_file(const char* path, const void* data, size_t n)
{
	FILE* f = std::fopen(path, "wb");
	if(!f)
		return false;
	const bool ok = std::fwrite(data, 1, n, f) == n;
	return (std::fclose(f) == 0) && ok;
}

std::string g_last_error;

} // namespace

extern "C" {

int swh_map_save_sqwm(const swh_map* map, const char* path)
{
	if(!map || !path)
		return SW_ERR_INVALID_ARGUMENT;
	std::vector<uint8_t> bytes;
	if(build_sqwm(*map, bytes) != 0)
		return (g_last_error = "entity strings above 65535 bytes"), SW_ERR_BUDGET;
	if(!write_file(path, bytes.data(), bytes.size()))
		return (g_last_error = std::string("cannot write ") + path), SW_ERR_INVALID_ARGUMENT;
	return SW_OK;
}

int swh_map_load_sqwm(const char* path, swh_map** out_map)
{
	if(!path || !out_map)
		return SW_ERR_INVALID_ARGUMENT;
	*out_map = nullptr;
	FILE* f = std::fopen(path, "rb");
	if(!f)
		return (g_last_error = std::string("cannot open ") + path), SW_ERR_INVALID_ARGUMENT;
	std::vector<uint8_t> bytes;
	uint8_t buf[1 << 16];
	size_t n;
	while((n = std::fread(buf, 1, sizeof(buf), f)) > 0)
		bytes.insert(bytes.end(), buf, buf + n);
	std::fclose(f);
	if(load_sqwm(bytes, out_map, g_last_error) != 0)
		return SW_ERR_INVALID_ARGUMENT;
	return SW_OK;
}

int swh_map_save_quake_map(const swh_map* map, const char* path)
{
	if(!map || !path)
		return SW_ERR_INVALID_ARGUMENT;
	const std::string s = build_quake_map(*map);
	if(!write_file(path, s.data(), s.size()))
		return (g_last_error = std::string("cannot write ") + path), SW_ERR_INVALID_ARGUMENT;
	return SW_OK;
}

const char* swh_last_error(void) { return g_last_error.c_str(); }

// The compression of lightmap_compression.rs, exposed for the tests.
void swh_compress_color(const float c[3], uint8_t out[4])
{
	const CompressedColor r = compress_color(c);
	std::memcpy(out, &r, 4);
}
void swh_decompress_color(const uint8_t c[4], float out[3])
{
	CompressedColor r;
	std::memcpy(&r, c, 4);
	decompress_color(r, out);
}
void swh_compress_vector(const float v[3], int8_t out_components[3], uint8_t* out_scale)
{
	const CompressedVector r = compress_vector(v3(v[0], v[1], v[2]));
	std::memcpy(out_components, r.components, 3);
	*out_scale = r.scale;
}

} // extern "C"
