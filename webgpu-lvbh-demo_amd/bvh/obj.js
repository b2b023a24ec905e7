'use strict'
// bvh/obj.js -- OBJ text -> the build's inputs, with parse-obj.js semantics.
//
// Reference: scripts/obj/parse-obj.js:3-87 (over the vendored OBJFile parser,
// scripts/obj/obj-file-parser.js).  Returns { NUM_TRIS, TRI_ARR, BOUNDS } as
// the reference does, so main.js's call `BVH.build(TRI_ARR, NUM_TRIS, BOUNDS)`
// (main.js:88) can take it unchanged:
//   * vertices of the first model only; OBJ (x, y, z) becomes (x, z, y)
//     (parse-obj.js:23-25);
//   * every face is fan-triangulated around its first vertex (:35-62);
//   * each vertex is padded with 3.1415 (:55-59);
//   * BOUNDS is the f64 min/max over the faces' vertices (:10-53), taken
//     before two 20,000-unit floor triangles at z_min + 0.01 are appended
//     (:65-78), so the floor lies outside BOUNDS.
// Face indices are 1-based; negative indices count back from the last vertex
// defined so far, and index 0 is an error (obj-file-parser.js:176-184).

const PAD = 3.1415
const FLOOR_SIZE = 10000.0

function parseObj (text) {
  const verts = []  // of the first model
  const faces = []
  // Models are created by 'o' lines, or by the first v/f line when none
  // exists yet (obj-file-parser.js:83-96); only the first one is used.
  let models = 0
  const lines = text.split('\n')
  for (const raw of lines) {
    const items = raw.replace(/\s+/g, ' ').trim().split(' ')
    if ((items[0] === 'v' || items[0] === 'f') && models === 0) models = 1
    switch (items[0]) {
      case 'o':
        models += 1
        break
      case 'v':
        if (models === 1) {
          verts.push([
            items.length >= 2 ? parseFloat(items[1]) : 0.0,
            items.length >= 3 ? parseFloat(items[2]) : 0.0,
            items.length >= 4 ? parseFloat(items[3]) : 0.0
          ])
        }
        break
      case 'f': {
        if (models !== 1) break
        const face = []
        for (let k = 1; k < items.length; k++) {
          if (items[k] === '') continue
          let idx = parseInt(items[k].split('/')[0])
          if (idx === 0) throw new Error('Faces uses invalid vertex index of 0')
          if (idx < 0) idx = verts.length + 1 + idx
          face.push(idx)
        }
        faces.push(face)
        break
      }
      default:
        break
    }
  }

  let xMin = 1e30; let yMin = 1e30; let zMin = 1e30
  let xMax = -1e30; let yMax = -1e30; let zMax = -1e30
  const tris = []
  const vertex = (i) => {
    const v = verts[i - 1]
    return [v[0], v[2], v[1]]  // y and z swapped
  }
  const grow = (p) => {
    xMin = Math.min(xMin, p[0]); yMin = Math.min(yMin, p[1]); zMin = Math.min(zMin, p[2])
    xMax = Math.max(xMax, p[0]); yMax = Math.max(yMax, p[1]); zMax = Math.max(zMax, p[2])
  }
  for (const face of faces) {
    if (face.length === 0) continue
    const vr = vertex(face[0])
    grow(vr)
    for (let y = 1; y < face.length - 1; y++) {
      const v1 = vertex(face[y])
      const v2 = vertex(face[y + 1])
      grow(v1)
      grow(v2)
      tris.push(vr[0], vr[1], vr[2], PAD, v1[0], v1[1], v1[2], PAD, v2[0], v2[1], v2[2], PAD)
    }
  }
  const floorHeight = zMin + 0.01
  const s = FLOOR_SIZE
  tris.push(
    -s, -s, floorHeight, PAD, s, -s, floorHeight, PAD, -s, s, floorHeight, PAD,
    s, -s, floorHeight, PAD, s, s, floorHeight, PAD, -s, s, floorHeight, PAD)
  return {
    NUM_TRIS: tris.length / 12,
    TRI_ARR: tris,
    BOUNDS: { min: [xMin, yMin, zMin], max: [xMax, yMax, zMax] }
  }
}

module.exports = { parseObj }
